/* idmx.h -- C ABI of libidmx.so, the MI355X (gfx950) Shan-Chen lattice
 * Boltzmann library behind idpy's ShanChenMultiPhase surface.
 *
 * Conventions
 *   - Every entry point returns int: 0 on success, a negative code on error
 *     (IDMX_E_*); idmx_last_error() returns the message of the calling
 *     thread's last error. This replaces the reference's only error channel,
 *     a RuntimeError raised when gcc fails (idpy/Utils/HostModule.py:226-230).
 *   - Device buffers are opaque addresses (void*) owned by the caller
 *     (allocated with idmx_alloc). Kernels borrow them. The reference passes
 *     device handles the same way, as opaque pointers
 *     (idpy/Utils/HostModule.py:273-285, idpy/Utils/CTypesTypes.py:63-66).
 *   - Work is enqueued on the context's HIP stream. idmx_d2h blocks until the
 *     copy is done; idmx_sync waits for all enqueued work.
 *   - A context, and the simulations and batches created on it, may be used
 *     by one host thread at a time: they share the context's stream and its
 *     reduction scratch buffers. Use one context per thread (or serialise the
 *     calls) for concurrent work; separate contexts are independent.
 *   - Lattice index g = x + y*L0 (+ z*L0*L1), x fastest
 *     (idpy/IdpyCode/IdpyUnroll.py:325-386). Populations are SoA,
 *     pop[g + q*V] (the reference's 'gpu' ordering, idpy/LBM/LBM.py:529-537).
 *     Density n, pseudopotential psi and velocity u (u[g + d*V]) are float64.
 *   - Populations are float64, or float32 "shifted" (stored value
 *     f_q - w_q*rho0) when the simulation is created with IDMX_PREC_F32S.
 */
#ifndef IDMX_H
#define IDMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IDMX_ABI_VERSION 1

enum {
  IDMX_OK = 0,
  IDMX_E_ARG = -1,      /* invalid argument */
  IDMX_E_HIP = -2,      /* HIP runtime error */
  IDMX_E_UNSUPPORTED = -3,
  IDMX_E_NOMEM = -4
};

/* Velocity sets (idpy/LBM/LBM.py:74-92). */
enum { IDMX_D2Q9 = 0, IDMX_D3Q19 = 1 };
/* Isotropic-gradient E-stencils, FStencils (idpy/LBM/LBM.py:94-185). */
enum { IDMX_E4 = 4, IDMX_E6 = 6, IDMX_E8 = 8 };
/* Pseudopotentials psi(n) (the psi_code/psi_sym kwargs,
 * idpy/LBM/MultiPhase.py:2120): */
enum {
  IDMX_PSI_EXP = 0,    /* 1 - exp(-n) */
  IDMX_PSI_EXPINV = 1, /* exp(-1/n)   */
  IDMX_PSI_LINEAR = 2, /* n           */
  IDMX_PSI_PROGRAM = 3 /* any psi_code: a stack program set with
                          idmx_sc_set_psi_program before the first step */
};
/* Collisions: BGK and MRT with the Guo forcing; IDMX_BGK_SCF is BGK with
 * the Shan-Chen forcing (equilibrium velocity u + F/(n omega), no Guo term,
 * u written as the mean of the pre- and post-collision velocities;
 * K_Collision_ShanChenMultiPhase[Walls], idpy/LBM/MultiPhaseKernels.py:
 * 132-304, used by MainLoopSimpleSC / MainLoopSimpleWallsSC).
 * IDMX_BGK_ISO (D2Q9, FP64 only) is BGK with the Guo forcing followed, on
 * every site, by the isotropy filter of the post-collision populations
 * (K_IsotropyFilter, idpy/LBM/LBMKernelsMeta.py:895-1080, used by
 * MainLoopIsoFilter, idpy/LBM/MultiPhase.py:1099-1223). Both run on one
 * fixed fused launch (not autotuned: the LDS-staged step on the x-shifted
 * storage, as launch variant 10, except 3D E6/E8), and on the two-pass step
 * with walls, a body force, a psi program or IDMX_RUN_TWO_PASS. */
enum { IDMX_BGK = 0, IDMX_MRT = 1, IDMX_BGK_SCF = 2, IDMX_BGK_ISO = 3 };
enum { IDMX_PREC_F64 = 0, IDMX_PREC_F32S = 1 };

/* Flags of idmx_sc_run. */
enum {
  IDMX_RUN_WRITE_FIELDS = 1, /* the last step also stores n, psi and the
                                Guo-shifted u of that step */
  IDMX_RUN_TWO_PASS = 2,     /* use the unfused two-kernel step (density/psi
                                kernel + force-collide-stream kernel), the
                                reference's own structure */
  IDMX_RUN_FUSED_WALLS = 4   /* with walls (FP64 BGK, E4): the fused step with
                                walls on any lattice size (by default it is
                                used from 2^19 sites up, the two-pass step
                                below; the last step of a run that writes the
                                fields is always the two-pass one) */
};

typedef struct idmx_ctx_s* idmx_ctx_t;
typedef struct idmx_sc_s* idmx_sc_t;

/* Description of one Shan-Chen simulation (the constructor kwargs of
 * ShanChenMultiPhase, idpy/LBM/MultiPhase.py:2096-2108). */
typedef struct {
  int xi;             /* IDMX_D2Q9 | IDMX_D3Q19                        */
  int estencil;       /* IDMX_E4 | IDMX_E6 | IDMX_E8                   */
  int dims[3];        /* L0, L1, L2 (L2 = 1 for D2Q9)                  */
  int psi_kind;       /* IDMX_PSI_*                                    */
  double SC_G;        /* coupling G                                    */
  double tau;         /* BGK relaxation time; OMEGA = 1/tau            */
  int collision;      /* IDMX_BGK | IDMX_MRT | IDMX_BGK_SCF | IDMX_BGK_ISO */
  int precision;      /* IDMX_PREC_F64 | IDMX_PREC_F32S                */
  double rho0;        /* population shift for IDMX_PREC_F32S           */
  /* MRT only (collision == IDMX_MRT), row-major float64 tables built on
   * the host from the relaxation matrix S (idpy/LBM/Collisions.py:96-160):
   *   mrt_A     Q x Q     M^-1 (1 - S) M
   *   mrt_beta  NB x Q    M^-1 S M v_k for the NB equilibrium basis vectors
   *   mrt_gamma NB x Q    M'^-1 (1 - S/2) M' v_k for the same vectors
   * with NB = 6 (D2Q9) or 10 (D3Q19), v_k = w∘{1, xi_a, xi_a xi_b (a<=b)}. */
  const double* mrt_A;
  const double* mrt_beta;
  const double* mrt_gamma;
  /* MRT with a diagonal S: its Q rates (then the three tables above may be
   * NULL); the kernel works in moment space with the Hermite bases compiled
   * in. */
  const double* mrt_diag;
} idmx_sc_desc;

/* ---- library / context ---------------------------------------------- */
int idmx_abi_version(void);
/* Copies the calling thread's last error message into buf (NUL-terminated). */
int idmx_last_error(char* buf, size_t len);
int idmx_device_count(int* count);
/* Creates a context on device `device` with its own HIP stream. Replaces the
 * CTypes/CUDA Tenet (idpy/CTypes/CTypes.py:59-83, idpy/CUDA/CUDA.py:46). */
int idmx_ctx_create(int device, idmx_ctx_t* out);
/* Destroys the context and its stream. Simulations and batches created on it
 * must be destroyed first; their handles are invalid afterwards. */
int idmx_ctx_destroy(idmx_ctx_t ctx);
int idmx_sync(idmx_ctx_t ctx);
/* Device name into buf (the 'device_name' entry of IdpyLoopProfile,
 * idpy/IdpyCode/IdpyCode.py:1913-2085). */
int idmx_device_name(idmx_ctx_t ctx, char* buf, size_t len);

/* ---- memory: IdpyMemory.Zeros/OnDevice, array.H2D/D2H
 *      (idpy/IdpyCode/IdpyMemory.py:760-865) ----------------------------- */
int idmx_alloc(idmx_ctx_t ctx, size_t bytes, void** dptr);
int idmx_free(idmx_ctx_t ctx, void* dptr);
int idmx_h2d(idmx_ctx_t ctx, void* dst, const void* src, size_t bytes);
int idmx_d2h(idmx_ctx_t ctx, void* dst, const void* src, size_t bytes);
int idmx_d2d(idmx_ctx_t ctx, void* dst, const void* src, size_t bytes);
int idmx_memset(idmx_ctx_t ctx, void* dst, int value, size_t bytes);

/* ---- reductions, FP64 (IdpyMemory.Sum/Max, IdpyMemory.py:865-889) ----- */
/* sum of n float64 values */
int idmx_sum_f64(idmx_ctx_t ctx, const double* x, int64_t n, double* out);
/* max of n float64 values (+inf if any is NaN) */
int idmx_max_f64(idmx_ctx_t ctx, const double* x, int64_t n, double* out);
/* max_i |a_i - b_i| (the |delta rho|_max steady-state check; +inf if any
 * difference is NaN) */
int idmx_max_abs_diff_f64(idmx_ctx_t ctx, const double* a, const double* b,
                          int64_t n, double* out);

/* idmx_max_abs_diff_f64 for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. It takes a HIP stream
 * (hipStream_t as void*, NULL for the null stream) of the current device
 * instead of a context and calls no hipMalloc, hipFree or allocating
 * routine: its partial results (at most 1024, 8 KiB) go to a static device
 * array, one copy per GPU, held until the process ends and looked up at the
 * first call on each device. Any n >= 0 is accepted (the reduction grid is
 * capped at 1024 workgroups). Calls that overlap on one GPU share that
 * array and race: serialise them. Blocks until the result is on the host. */
int idmx_max_abs_diff_f64_noalloc(void* stream, const double* a, const double* b,
                                  int64_t n, double* out);

/* K_CenterOfMass + IdpyMemory.Sum (idpy/LBM/MultiPhase.py:108-137) over a
 * float64 density field of L0 x L1 x L2 sites (x fastest; L2 = 1 in 2D):
 * out[0] = sum n, out[1 + d] = sum n x_d / sum n, coordinates not
 * unwrapped across the periodic boundary. */
int idmx_center_of_mass_f64(idmx_ctx_t ctx, const double* n, int L0, int L1,
                            int L2, double* out);

/* dst[i] = src[i] for n float64 values with a plain one-element-per-thread
 * kernel; *ms = its duration. Calibrates achievable HBM bandwidth and the
 * FETCH_SIZE / WRITE_SIZE counters (2 n 8 bytes moved). */
int idmx_stream_copy_f64(idmx_ctx_t ctx, double* dst, const double* src,
                         int64_t n, double* ms);

/* K_CheckUMeta followed by IdpyMemory.Sum/Max
 * (idpy/LBM/MultiPhaseLoopChecks.py:96-163): *sum_delta = sum_g sum_d
 * |u - old_u|, *max_u2 = max_g sum_d u^2; old_u is overwritten with u.
 * u and old_u are SoA float64 with D components. */
int idmx_check_u(idmx_ctx_t ctx, const double* u, double* old_u, int64_t V,
                 int D, double* sum_delta, double* max_u2);

/* ---- initial interfaces: n from the lattice position, u = 0 -----------
 * n and u are float64 device arrays of V and D*V values on a D-dimensional
 * lattice dims[0..D-1] (x fastest), centred on dim_center = dims/2
 * (idpy/LBM/LBM.py:422-423); full_flag != 0 puts the liquid n_l inside
 * (delta_n = n_l - n_g), 0 the gas. */
/* K_InitRadialInterface: n = (n_g+n_l)/2 - delta_n/2 tanh((r - R)/W), r the
 * distance to the centre (legacy class idpy/LBM/LBM.py:1786-1826; the
 * Meta class's kernel, idpy/LBM/MultiPhaseKernels.py:686-727, is W = 1). */
int idmx_init_radial(idmx_ctx_t ctx, double* n, double* u, int D, const int* dims,
                     double n_g, double n_l, double R, double W, int full_flag);
/* K_InitFlatInterface (idpy/LBM/LBM.py:1745-1784) with
 * F_NFlatProfilePeriodic (idpy/LBM/MultiPhaseKernels.py:55-74): a periodic
 * slab of width `width` centred along `direction`. */
int idmx_init_flat(idmx_ctx_t ctx, double* n, double* u, int D, const int* dims,
                   double n_g, double n_l, double width, int direction, int full_flag);
/* K_InitCylinderInterface (idpy/LBM/LBM.py:1828-1874): the radial profile
 * (W = 1) of the distance in the plane orthogonal to `direction`. */
int idmx_init_cylinder(idmx_ctx_t ctx, double* n, double* u, int D, const int* dims,
                       double n_g, double n_l, double R, int direction, int full_flag);
/* K_InitSingleFlatInterface (idpy/LBM/MultiPhaseKernels.py:650-684):
 * n = n_g + delta_n (tanh(x_dir - c_dir) + 1)/2. */
int idmx_init_single_flat(idmx_ctx_t ctx, double* n, double* u, int D, const int* dims,
                          double n_g, double n_l, int direction, int full_flag);

/* ---- Shan-Chen simulation -------------------------------------------- */
int idmx_sc_create(idmx_ctx_t ctx, const idmx_sc_desc* desc, idmx_sc_t* out);
int idmx_sc_destroy(idmx_sc_t sim);
/* Bytes of one population buffer (Q * V * sizeof(population)). */
int idmx_sc_pop_bytes(idmx_sc_t sim, size_t* bytes);

/* K_InitPopulations: pop = f_eq(n, u), 2nd-order Hermite equilibrium
 * (idpy/LBM/LBM.py:1700-1743; K_InitPopulationsMeta
 * idpy/LBM/LBMKernelsMeta.py:282-387). u is SoA float64. */
int idmx_sc_init_pop(idmx_sc_t sim, void* pop, const double* n,
                     const double* u);

/* K_InitPopulationsMeta as the current class generates it: f_eq with the
 * coefficients it prints, 15 significant digits (str(expand(f_eq).evalf()),
 * idpy/LBM/Equilibria.py:202-210, idpy/LBM/LBMKernelsMeta.py:282-390).
 * coeffs is a host array of NB x Q doubles (NB = 6 in 2D, 10 in 3D):
 * f_q = sum_k coeffs[k*Q + q] c_k with the equilibrium coordinates
 * c_0 = n - 1.5 n u.u, c_a = 3 n u_a, c_aa = 4.5 n u_a^2, c_ab = 9 n u_a u_b
 * (basis order 2D {1, x, y, xx, yy, xy}, 3D {1, x, y, z, xx, yy, zz, xy,
 * xz, yz}); the compiled weights give idmx_sc_init_pop. u is SoA float64. */
int idmx_sc_init_pop_coeffs(idmx_sc_t sim, void* pop, const double* n,
                            const double* u, const double* coeffs);

/* K_ComputeDensityPsiMeta(n, psi, pop): n = sum_q f_q, psi = psi(n)
 * (idpy/LBM/MultiPhaseKernelsMeta.py:107-167). */
int idmx_sc_density_psi(idmx_sc_t sim, double* n, double* psi,
                        const void* pop);

/* K_ForceCollideStreamSCMPMeta(pop_swap, pop, n, psi) and
 * K_ForceCollideStreamSCMP_MRTMeta: Shan-Chen force from psi, Guo-shifted
 * velocity, BGK or MRT collision with Guo forcing, push streaming into
 * pop_swap (idpy/LBM/MultiPhaseKernelsMeta.py:169-345, 891-1075). */
int idmx_sc_force_collide_stream(idmx_sc_t sim, void* pop_swap,
                                 const void* pop, const double* n,
                                 const double* psi);

/* The MainLoop time loop (idpy/LBM/MultiPhase.py:1015-1097, legacy
 * idpy/LBM/LBM.py:806-894): nsteps fused steps, each reading *pop and
 * writing *pop_swap, then swapping the two addresses (M_SwapPop,
 * idpy/LBM/LBMKernels.py:528-540). On return *pop holds the current state.
 * With IDMX_RUN_WRITE_FIELDS, n, psi and u (Guo-shifted, SoA) of the last
 * step are stored, which is what the reference's n/psi (both classes) and u
 * (legacy class) arrays hold after MainLoop. n, psi, u may be NULL without
 * that flag. */
int idmx_sc_run(idmx_sc_t sim, void** pop, void** pop_swap, double* n,
                double* psi, double* u, int nsteps, int flags);

/* idmx_sc_run with per-step kernel times, as IdpyLoopProfile returns them
 * (idpy/IdpyCode/IdpyCode.py:1913-2085; MainLoop(profiling=True),
 * idpy/LBM/MultiPhase.py:1048-1097): step_ms holds 2 * nsteps floats,
 * [2k] the density/psi kernel of step k and [2k+1] its force-collide-stream
 * kernel on the two-pass step, or [2k] = 0 and [2k+1] the fused kernel;
 * milliseconds from HIP events on the library stream. The pull-form
 * conversions before and after the steps are not in any step's time. */
int idmx_sc_run_profiled(idmx_sc_t sim, void** pop, void** pop_swap, double* n,
                         double* psi, double* u, int nsteps, int flags,
                         float* step_ms);

/* ComputeMomentsMeta: n, psi and u = sum xi f / n + F/(2n)
 * (idpy/LBM/MultiPhase.py:1565-1662,
 *  K_ComputeVelocityAfterForceSCMPMeta MultiPhaseKernelsMeta.py:1077-1219). */
int idmx_sc_moments(idmx_sc_t sim, const void* pop, double* n, double* psi,
                    double* u);

/* K_ComputeMoments of the legacy class: n = sum_q f_q and the unshifted
 * u = sum_q xi_q f_q / n (idpy/LBM/LBM.py:1667-1698). */
int idmx_sc_density_velocity(idmx_sc_t sim, const void* pop, double* n,
                             double* u);

/* Total mass sum_g n(g) and momentum sum_g sum_q xi_q f_q(g) of a
 * population buffer, FP64 accumulation; mom has D entries. */
int idmx_sc_mass_momentum(idmx_sc_t sim, const void* pop, double* mass,
                          double* mom);

/* Converts between the device population layout and a float64 SoA host
 * array (for FP32S simulations this adds/removes the w_q*rho0 shift). */
int idmx_sc_pop_to_f64(idmx_sc_t sim, const void* pop, double* dst_dev);
int idmx_sc_pop_from_f64(idmx_sc_t sim, void* pop, const double* src_dev);

/* Launch shape of the fused step: tile tx x ty (2D: tx = 64, 128 or 256
 * sites of one row, ty = 1; 3D: 32x8, 64x4, 128x2, 256x2), H planes marched
 * per block, and the variant `hist`: 1 holds the tile's populations in
 * registers between their read and their collision, 0 re-reads them, 2
 * re-reads them and prefetches the next plane before the collision, 3 and 4
 * are 0 and 2 compiled for at least 3 waves per SIMD, 5 and 6 are 1 and 0
 * run on the pull form (post-collision values stored at their own site,
 * read from the source site; idmx_sc_run converts to that form before the
 * steps and back after them, so pop always holds streamed populations
 * between calls), 7 and 8 are 0 and 6 with two lead planes loaded per
 * barrier, 9 holds them in registers and copies each lead plane (tile and
 * halo) into LDS by DMA while the previous plane collides (refused with
 * IDMX_E_ARG for tiles whose staging buffer exceeds the 160 KiB of LDS of a
 * CU, e.g. FP64 256x2), 10 is 9 with the populations kept, between the
 * steps of one idmx_sc_run, in an x-shifted storage (population q of site x
 * at x - c_qx, so every push store is line-aligned; the run's first step
 * reads the streamed storage and its last writes it, so pop holds streamed
 * populations between calls, and the results equal variant 9's bit for
 * bit). Chosen at creation;
 * the autotuner (idmx.tune) overrides it. No reference counterpart: the
 * reference fixes block = 128 threads (idpy/LBM/MultiPhase.py:2072-2085). */
int idmx_sc_set_launch(idmx_sc_t sim, int tx, int ty, int H, int hist);

/* The pseudopotential of a simulation created with IDMX_PSI_PROGRAM: the
 * C expression psi_code that the reference compiles into
 * K_ComputeDensityPsiMeta (idpy/LBM/MultiPhaseKernelsMeta.py:107-167,
 * MultiPhase.py:2120), compiled on the host into a stack program of nops
 * op codes (<= 48) and nconsts FP64 constants (<= 16), evaluated in FP64 at
 * every site: 0 push n, 1 push the next constant, 2 add, 3 sub, 4 mul,
 * 5 div, 6 negate, 7 exp, 8 log, 9 sqrt, 10 pow, 11 fabs, 12 tanh, 13 sinh,
 * 14 cosh, 15 sin, 16 cos, 17 fmax, 18 fmin, 19 round to float; at most 8
 * values on the stack and exactly one at the end (checked, IDMX_E_ARG
 * otherwise). The program is copied to the device. */
int idmx_sc_set_psi_program(idmx_sc_t sim, const unsigned char* ops, int nops,
                            const double* consts, int nconsts);

/* Walls: `walls` is a device array of V unsigned chars (the reference's
 * FlagType), 1 = fluid, 0 = solid; NULL removes them. The simulation borrows
 * it. With walls, idmx_sc_run steps as MainLoopWalls
 * (idpy/LBM/MultiPhase.py:1226-1309): K_ComputeDensityPsiWallsMeta on the
 * fluid sites, then K_ForceCollideStreamWallsSCMPMeta (fluid sites, push
 * with half-way bounce-back at solid sites;
 * idpy/LBM/MultiPhaseKernelsMeta.py:42-104, 524-716), and needs the n and psi
 * buffers: psi at solid sites keeps what the caller stored (GetWalls,
 * MultiPhase.py:157-161). idmx_sc_density_psi, idmx_sc_force_collide_stream
 * and idmx_sc_moments honour the walls the same way. */
int idmx_sc_set_walls(idmx_sc_t sim, const void* walls);

/* Body force (fx, fy, fz; fz ignored in 2D) added to the Shan-Chen force at
 * every updated site, in the velocity shift and the Guo term: the gravity of
 * MainLoopSimpleWallsGravity (F[1] += g,
 * idpy/LBM/MultiPhaseKernels.py:474-560, idpy/LBM/MultiPhase.py:888-1010).
 * Zero by default; while non-zero, idmx_sc_run uses the two-pass step (n and
 * psi buffers required) and idmx_sc_moments includes it in u. */
int idmx_sc_set_body_force(idmx_sc_t sim, double fx, double fy, double fz);
int idmx_sc_get_launch(idmx_sc_t sim, int* tx, int* ty, int* H, int* hist);

/* Milliseconds of the last idmx_sc_run, measured with HIP events recorded
 * on the simulation's stream around its kernels. */
int idmx_sc_last_run_ms(idmx_sc_t sim, double* ms);

/* ---- batched sweep over many D2Q9 domains ----------------------------
 * No single reference counterpart: the reference runs one
 * ShanChenMultiPhase per (G, R, L) point and compiles every domain size into
 * its own kernel (sizes are #defines, idpy/LBM/LBM.py:542-569). Here one
 * launch advances all live domains of a sweep (BGK, tau per domain, periodic,
 * centred or arbitrary initial densities), with populations packed in
 * library-owned HBM buffers. */
typedef struct idmx_batch_s* idmx_batch_t;
typedef struct {
  int L0, L1;        /* domain size                                      */
  double SC_G;       /* coupling                                         */
  double tau;        /* BGK relaxation time                              */
  double rho0;       /* population shift (IDMX_PREC_F32S)                */
} idmx_batch_domain;

int idmx_batch_create(idmx_ctx_t ctx, const idmx_batch_domain* doms, int ndom,
                      int precision, int estencil, int psi_kind,
                      idmx_batch_t* out);
int idmx_batch_destroy(idmx_batch_t b);
/* Work decomposition. kind 0 (march): items are x-chunks of tx sites
 * (64, 128, 256) by runs of H rows, each marched row by row with the psi
 * ring in LDS (large domains). kind 1 (tile, the default 32 x 8): items are
 * tx x H tiles (64x4, 32x8, 16x16, 32x4), one site per thread, halo psi in
 * LDS (small domains: no dependent row loads). kind 2 / 3: kind 0 / 1 run on
 * the pull form (post-collision values at their own site, read from the
 * source site: aligned stores); idmx_batch_run converts every live domain to
 * that form before its steps and back after them. kind 4 (E4 only): tx x H
 * tiles (32x8, 16x16, 32x16) advanced two steps per launch, the first
 * step's populations kept in LDS for the tile plus a 2-site ring (half the
 * HBM bytes and half the launches per step; an odd step left over runs as
 * kind 1); the same results as kind 1, bit for bit. kind 5: kind 4 with
 * the populations kept, between the launches of idmx_batch_run's captured
 * graph, in an x-shifted storage (population q of site x at x - c_qx, so
 * every push store covers the tile's own x range); each replay reads and
 * writes the streamed storage at its ends, and the results equal kind 4's
 * bit for bit. grid = blocks of a step
 * launch (0 keeps the current grid); a step launches at most one block per
 * live work item, and each XCD takes a contiguous run of items. Picked by
 * the autotuner (idmx.tune). */
int idmx_batch_set_launch(idmx_batch_t b, int kind, int tx, int H, int grid);
int idmx_batch_get_launch(idmx_batch_t b, int* kind, int* tx, int* H, int* grid,
                          int* n_items);
/* active[d] != 0 keeps domain d in the work list. */
int idmx_batch_set_active(idmx_batch_t b, const int* active);
/* f = f_eq(n, 0) for domain dom from a host float64 density (x fastest);
 * also resets the domain's reference for idmx_batch_check. */
int idmx_batch_set_density(idmx_batch_t b, int dom, const double* n_host);
int idmx_batch_get_density(idmx_batch_t b, int dom, double* n_host);
/* Advances every live domain by nsteps (graph replays of 100 steps where
 * nsteps allows). */
int idmx_batch_run(idmx_batch_t b, int nsteps);
/* dmax[d] = max_x |n(x) - n_prev(x)| over live domain d in FP64, where
 * n_prev is the density at the previous check (or set_density); 0 for
 * inactive domains. */
int idmx_batch_check(idmx_batch_t b, double* dmax);
/* Copies the state of every domain from src to dst (same domains; FP32
 * shifted <-> FP64 conversion included). flags = IDMX_BATCH_COPY_REST
 * writes the rest equilibrium f = w n(src) instead (zero momentum: the
 * FP32 -> FP64 hand-over of a sweep, where FP32 rounding has left a small
 * net momentum that FP64 would conserve). */
#define IDMX_BATCH_COPY_REST 1
int idmx_batch_copy_state(idmx_batch_t dst, idmx_batch_t src, int flags);
/* Projects the density of domain dom (-1: every live domain) onto the
 * fields symmetric about its centre (L0/2, L1/2): the average over the
 * mirror images in x and y and, for square domains, the transpose. With
 * target_mass (ndom values, or NULL to keep the mass) a uniform shift then
 * sets the domain's total mass to target_mass[d]. The populations become
 * the rest equilibrium of the result and it becomes the check reference.
 * For centred droplets, whose steady state has this symmetry: FP32 rounding
 * or an approximate start breaks it (an off-centre droplet drifts towards a
 * symmetric lattice position over ~1e5 steps), and FP32 storage drifts the
 * mass by ~1e-5 relative over a few thousand steps. */
int idmx_batch_symmetrize(idmx_batch_t b, int dom, const double* target_mass);
/* Rewrites the populations of domain dom (-1: every live domain) as the
 * fluid at rest under the Guo forcing, from its check-reference density n
 * (the density the last set_density / symmetrize / check left there):
 * f = f_eq(n, -F / (2n)) with F = -G psi(x) sum_k W_k e_k psi(x + e_k) the
 * domain's Shan-Chen force, so the collision velocity (sum_q xi_q f_q +
 * F/2) / n is zero. The reference's InitPopulations at u = 0 (f = w n, as
 * set_density writes) instead starts each interface at the velocity
 * F / (2n). For warm starts: a start at the steady density with f = w n
 * needs ~95 % of a cold start's steps to |dn| < 1e-10, with these
 * populations a few hundred. No reference counterpart. */
int idmx_batch_force_rest(idmx_batch_t b, int dom);
/* Copies domain dom's populations to the host as float64 SoA (9 x V, the
 * w_q rho0 shift of FP32 storage added back). */
int idmx_batch_get_populations(idmx_batch_t b, int dom, double* f_host);
int idmx_batch_last_run_ms(idmx_batch_t b, double* ms);

/* ---- warm-start network kernels ---------------------------------------
 * No reference counterpart (idpy has no learned component). Tensors are NHWC
 * on the device; activations and conv weights in bf16 (uint16 bit patterns),
 * master weights, gradients and losses in float32. A conv weight is
 * [Cout][9*Cin] with k = (3*ky + kx)*Cin + ci. Channel counts that are
 * multiples of 4 take the vector-load path. */
/* out = act(conv3x3(in, wt) + bias), zero padding 1; out_f32 (may be NULL)
 * receives the FP32 pre-activation; relu != 0 applies max(0, .). */
int idmx_nn_conv3x3_fwd(idmx_ctx_t ctx, const void* in, const void* wt,
                        const float* bias, void* out, float* out_f32, int N,
                        int H, int W, int Cin, int Cout, int relu);
/* dx = conv3x3^T(dy, w); w is the FP32 master weight, wd_scratch holds
 * 9*Cin*Cout bf16 values. */
int idmx_nn_conv3x3_bwd_data(idmx_ctx_t ctx, const void* dy, const float* w,
                             void* wd_scratch, void* dx, int N, int H, int W,
                             int Cin, int Cout);
/* dw += sum over pixels of dy (x) im2col(in); db += sum of dy (db may be
 * NULL). */
int idmx_nn_conv3x3_bwd_weight(idmx_ctx_t ctx, const void* in, const void* dy,
                               float* dw, float* db, int N, int H, int W,
                               int Cin, int Cout);
int idmx_nn_relu_bwd(idmx_ctx_t ctx, const void* y, const void* dy, void* dx,
                     int64_t n);
/* 2x2 average pool (bwd = 0: x is N,H,W,C -> y N,H/2,W/2,C; bwd = 1: x is the
 * pooled gradient, y the N,H,W,C gradient). */
int idmx_nn_pool2(idmx_ctx_t ctx, const void* x, void* y, int N, int H, int W,
                  int C, int bwd);
/* bwd = 0: nearest 2x upsample of x (N,H,W,C) into channels [c0, c0+C) of
 * y (N,2H,2W,Cy). bwd = 1: x is a gradient of shape (N,2H,2W,Cy); y
 * (N,H,W,C) receives the sum over each 2x2 block of its channels
 * [c0, c0+C). */
int idmx_nn_up2(idmx_ctx_t ctx, const void* x, void* y, int N, int H, int W,
                int C, int Cy, int c0, int bwd);
int idmx_nn_copy_channels(idmx_ctx_t ctx, const void* x, void* y, int64_t P,
                          int Cx, int cx0, int Cy, int cy0, int C);
int idmx_nn_add(idmx_ctx_t ctx, const void* a, const void* b, void* y,
                int64_t n);
/* loss = mean (pred - tgt)^2, grad = 2 (pred - tgt) / n in bf16. */
int idmx_nn_mse(idmx_ctx_t ctx, const float* pred, const float* tgt,
                void* grad, int64_t n, double* loss);
/* Adam step t >= 1 on FP32 w with gradient g (zeroed afterwards); refreshes
 * the bf16 copy wb (may be NULL). */
int idmx_nn_adam(idmx_ctx_t ctx, float* w, float* g, float* m, float* v,
                 void* wb, int64_t n, float lr, float b1, float b2, float eps,
                 int t);
/* float32 <-> bf16 */
int idmx_nn_cast(idmx_ctx_t ctx, const void* x, void* y, int64_t n,
                 int to_bf16);
/* Bilinear resize of an S x S float32 image to an L x L float64 lattice
 * density, plus a constant. */
int idmx_nn_resize_to_lattice(idmx_ctx_t ctx, const float* img, int S,
                              double* n, int L, double add);

#ifdef __cplusplus
}
#endif
#endif /* IDMX_H */
