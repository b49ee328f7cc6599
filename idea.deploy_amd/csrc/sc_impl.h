// sc_impl.h -- host-side launch logic for one (velocity set, E-stencil,
// precision, collision) combination. Included by sc_d2q9.hip / sc_d3q19.hip,
// which instantiate the combinations of their velocity set.
#pragma once
#include <algorithm>

#include "internal.h"

namespace idmx {

inline unsigned grid_for(int64_t V) {
  const int64_t b = (V + 255) / 256;
  return (unsigned)std::min<int64_t>(b, 65536);
}

// Launch variants that run on the pull form of the populations.
inline bool pull_form(int variant) { return variant == 5 || variant == 6 || variant == 8; }

template <class S, class ES, class T, bool SHIFT, int COLL>
struct Impl {
  static Phys<T> phys(const idmx_sc_s* s) {
    Phys<T> P;
    P.G = (T)s->desc.SC_G;
    P.omega = (T)(1.0 / s->desc.tau);
    P.rho0 = SHIFT ? (T)s->desc.rho0 : T(0);
    P.psi_kind = s->desc.psi_kind;
    const T* m = (const T*)s->mrt_dev;
    constexpr int NB = NBasis<S::D>::value;
    P.mrtA = m;
    P.mrtB = m ? m + S::Q * S::Q : nullptr;
    P.mrtG = m ? m + S::Q * S::Q + NB * S::Q : nullptr;
    P.prog = s->psi_prog_dev;
    return P;
  }

  // K_ComputeDensityPsiMeta; a psi_code program has its own instantiation
  static void launch_density_psi(idmx_sc_s* s, dim3 grid, dim3 block, unsigned shm,
                                 hipStream_t st, const T* pop, Geom g, Phys<T> P, double* n,
                                 double* psi, const uint8_t* walls) {
    if (psi && s->desc.psi_kind == IDMX_PSI_PROGRAM)
      hipLaunchKernelGGL((k_density_psi<S, T, SHIFT, true>), grid, block, shm, st, pop, g, P, n,
                         psi, walls);
    else
      hipLaunchKernelGGL((k_density_psi<S, T, SHIFT>), grid, block, shm, st, pop, g, P, n, psi,
                         walls);
  }

  template <int TX, int TY, bool WF, bool HIST, bool PF = false, bool O3 = false,
            bool PULL = false, bool P2 = false>
  static hipError_t launch_fused(idmx_sc_s* s, const T* pop, T* swap, double* n,
                                 double* psi, double* u) {
    const Geom& g = s->geo;
    const int ntx = (g.L0 + TX - 1) / TX;
    const int nty = S::D == 3 ? (g.L1 + TY - 1) / TY : 1;
    const int Lm = S::D == 3 ? g.L2 : g.L1;
    const int H = std::min(s->H, Lm);
    const int nseg = (Lm + H - 1) / H;
    const int ntiles = ntx * nty * nseg;
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
    if constexpr (O3)
      hipLaunchKernelGGL((k_fused_o3<S, ES, T, SHIFT, COLL, WF, TX, TY, HIST, PF>), dim3(grid),
                         dim3(TX * TY), 0, s->ctx->stream, pop, swap, g, phys(s), H, ntx, nty,
                         ntiles, n, psi, u);
    else
      hipLaunchKernelGGL((k_fused<S, ES, T, SHIFT, COLL, WF, TX, TY, HIST, PF, PULL, P2>),
                         dim3(grid), dim3(TX * TY), 0, s->ctx->stream, pop, swap, g, phys(s), H,
                         ntx, nty, ntiles, n, psi, u);
    return hipGetLastError();
  }

  // Lead plane staged in LDS by DMA (variant 9); tiles whose staging buffer
  // does not fit the LDS of a CU are refused.
  template <int TX, int TY, bool XSH = false>
  static hipError_t launch_dma(idmx_sc_s* s, const T* pop, T* swap) {
    if constexpr (DmaShape<S, ES, T, TX, TY, COLL>::fits) {
      const Geom& g = s->geo;
      const int ntx = (g.L0 + TX - 1) / TX;
      const int nty = S::D == 3 ? (g.L1 + TY - 1) / TY : 1;
      const int Lm = S::D == 3 ? g.L2 : g.L1;
      const int H = std::min(s->H, Lm);
      const int nseg = (Lm + H - 1) / H;
      const int ntiles = ntx * nty * nseg;
      const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
      hipLaunchKernelGGL((k_fused_dma<S, ES, T, SHIFT, COLL, TX, TY, false, XSH>), dim3(grid),
                         dim3(TX * TY), 0, s->ctx->stream, pop, swap, g, phys(s), H, ntx, nty,
                         ntiles, nullptr, nullptr, T(0), T(0), T(0), XSH ? s->xsh : 0);
      return hipGetLastError();
    } else {
      return hipErrorInvalidConfiguration;
    }
  }

  // The fused step with walls and a body force (k_fused_dma<..., WALLS>):
  // 3D 64x4 tiles (32x8 where the staging does not fit), 2D rows of 256 (64
  // on narrow lattices); FP64 BGK only (kFusedWalls).
  template <int TX, int TY>
  static hipError_t launch_dma_walls(idmx_sc_s* s, const T* pop, T* swap, const double* psi) {
    // on the x-shifted storage between the steps of a run (s->xsh)
    if constexpr (DmaShape<S, ES, T, TX, TY, COLL, true>::fits && !SHIFT && COLL == 0) {
      const Geom& g = s->geo;
      const int ntx = (g.L0 + TX - 1) / TX;
      const int nty = S::D == 3 ? (g.L1 + TY - 1) / TY : 1;
      const int Lm = S::D == 3 ? g.L2 : g.L1;
      const int H = std::min(s->H, Lm);
      const int nseg = (Lm + H - 1) / H;
      const int ntiles = ntx * nty * nseg;
      const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
      hipLaunchKernelGGL((k_fused_dma<S, ES, T, SHIFT, COLL, TX, TY, true, true>), dim3(grid),
                         dim3(TX * TY), 0, s->ctx->stream, pop, swap, g, phys(s), H, ntx, nty,
                         ntiles, psi, (const uint64_t*)s->wmask_dev, (T)s->bforce[0],
                         (T)s->bforce[1], (T)s->bforce[2], s->xsh);
      return hipGetLastError();
    } else {
      return hipErrorInvalidConfiguration;
    }
  }
  // BGK with the E4 force only (the reference's wall loops are BGK; the MRT
  // step with the two extra staged planes, and the 3D step with a reach-2
  // stencil, need 256 VGPRs, one wave per SIMD; they stay on the two-pass
  // step)
  static constexpr bool kFusedWalls =
      !SHIFT && COLL == 0 && ES::R == 1 &&
      (S::D == 2 ? DmaShape<S, ES, T, 256, 1, COLL, true>::fits &&
                       DmaShape<S, ES, T, 64, 1, COLL, true>::fits
                 : DmaShape<S, ES, T, 32, 8, COLL, true>::fits);
  static hipError_t fused_walls(idmx_sc_s* s, const T* pop, T* swap, const double* psi) {
    if constexpr (S::D == 2) {
      if (s->geo.L0 >= 256) return launch_dma_walls<256, 1>(s, pop, swap, psi);
      return launch_dma_walls<64, 1>(s, pop, swap, psi);
    } else {
      if constexpr (DmaShape<S, ES, T, 64, 4, COLL, true>::fits)
        return launch_dma_walls<64, 4>(s, pop, swap, psi);
      else
        return launch_dma_walls<32, 8>(s, pop, swap, psi);
    }
  }

  // variant: 1 populations held in registers, 0 re-read, 2 re-read with
  // prefetch of the next lead plane, 3 re-read with >= 3 waves/SIMD,
  // 4 re-read with prefetch and >= 3 waves/SIMD; 5 and 6 are 1 and 0 on the
  // pull form (run() converts to and from it around the steps); 7 and 8
  // are 0 and 6 with two lead planes per barrier; 9 stages the lead plane in
  // LDS by DMA (k_fused_dma); 10 is 9 with the populations kept in the
  // x-shifted storage between the steps of a run
  template <int TX, int TY>
  static hipError_t variant(idmx_sc_s* s, const T* pop, T* swap) {
    switch (s->hist) {
      case 9: return launch_dma<TX, TY>(s, pop, swap);
      case 10: return launch_dma<TX, TY, true>(s, pop, swap);
      case 7: return launch_fused<TX, TY, false, false, false, false, false, true>(
          s, pop, swap, nullptr, nullptr, nullptr);
      case 8: return launch_fused<TX, TY, false, false, false, false, true, true>(
          s, pop, swap, nullptr, nullptr, nullptr);
      case 5: return launch_fused<TX, TY, false, true, false, false, true>(s, pop, swap, nullptr,
                                                                          nullptr, nullptr);
      case 6: return launch_fused<TX, TY, false, false, false, false, true>(s, pop, swap, nullptr,
                                                                           nullptr, nullptr);
      case 0: return launch_fused<TX, TY, false, false>(s, pop, swap, nullptr, nullptr, nullptr);
      case 2: return launch_fused<TX, TY, false, false, true>(s, pop, swap, nullptr, nullptr,
                                                              nullptr);
      case 3: return launch_fused<TX, TY, false, false, false, true>(s, pop, swap, nullptr,
                                                                     nullptr, nullptr);
      case 4: return launch_fused<TX, TY, false, false, true, true>(s, pop, swap, nullptr,
                                                                    nullptr, nullptr);
      default: return launch_fused<TX, TY, false, true>(s, pop, swap, nullptr, nullptr, nullptr);
    }
  }

  // The Shan-Chen forcing (COLL 3) and the isotropy filter (COLL 4) on a
  // fixed fused launch (push form, populations held in registers), not on
  // the autotuner's variants: rows of 256 sites on lattices at least that
  // wide (256 x 2 tiles in 3D), else 64 x 1 (64 x 4); the step that writes
  // n / psi / u uses 64 x 1 (32 x 8).
  template <bool WF>
  static hipError_t fused_fixed(idmx_sc_s* s, const T* pop, T* swap, double* n, double* psi,
                                double* u) {
    if constexpr (WF) {
      if constexpr (S::D == 2) return launch_fused<64, 1, true, true>(s, pop, swap, n, psi, u);
      else return launch_fused<32, 8, true, true>(s, pop, swap, n, psi, u);
    } else {
      // the LDS-staged step on the x-shifted storage (s->xsh) where it
      // applies, else the populations held in registers
      if constexpr (S::D == 2) {
        if (s->geo.L0 >= 256) {
          if constexpr (kFixedDma2) return launch_dma<256, 1, true>(s, pop, swap);
          return launch_fused<256, 1, false, true>(s, pop, swap, nullptr, nullptr, nullptr);
        }
        if constexpr (kFixedDma2) return launch_dma<64, 1, true>(s, pop, swap);
        return launch_fused<64, 1, false, true>(s, pop, swap, nullptr, nullptr, nullptr);
      } else {
        if constexpr (kFixedDma3) return launch_dma<64, 4, true>(s, pop, swap);
        if (s->geo.L0 >= 256)
          return launch_fused<256, 2, false, true>(s, pop, swap, nullptr, nullptr, nullptr);
        return launch_fused<64, 4, false, true>(s, pop, swap, nullptr, nullptr, nullptr);
      }
    }
  }
  // whether fused_fixed<false> runs the LDS-staged step: 2D rows of 256 and
  // 64 sites, 3D 64x4 tiles with the reach-1 stencil (a reach-2 staging
  // leaves one workgroup per CU)
  static constexpr bool kFixedDma2 = S::D == 2 && DmaShape<S, ES, T, 256, 1, COLL>::fits &&
                                     DmaShape<S, ES, T, 64, 1, COLL>::fits;
  static constexpr bool kFixedDma3 =
      S::D == 3 && ES::R == 1 && DmaShape<S, ES, T, 64, 4, COLL>::fits;

  // Tile shapes the autotuner chooses from; the step that writes n/psi/u at
  // a checkpoint uses one fixed shape.
  template <bool WF>
  static hipError_t fused(idmx_sc_s* s, const T* pop, T* swap, double* n,
                          double* psi, double* u) {
    if constexpr (WF) {
      if (pull_form(s->hist)) {
        if constexpr (S::D == 2)
          return launch_fused<64, 1, true, true, false, false, true>(s, pop, swap, n, psi, u);
        else
          return launch_fused<32, 8, true, true, false, false, true>(s, pop, swap, n, psi, u);
      }
      if constexpr (S::D == 2) return launch_fused<64, 1, true, true>(s, pop, swap, n, psi, u);
      else return launch_fused<32, 8, true, true>(s, pop, swap, n, psi, u);
    } else {
      if constexpr (S::D == 2) {
        switch (s->tx) {
          case 64: return variant<64, 1>(s, pop, swap);
          case 128: return variant<128, 1>(s, pop, swap);
          default: return variant<256, 1>(s, pop, swap);
        }
      } else {
        switch (s->tx * 100 + s->ty) {
          case 6404: return variant<64, 4>(s, pop, swap);
          case 12802: return variant<128, 2>(s, pop, swap);
          case 25602: return variant<256, 2>(s, pop, swap);
          default: return variant<32, 8>(s, pop, swap);
        }
      }
    }
  }

  // mid: event recorded between the two kernels (per-step profiling), or null
  static hipError_t two_pass(idmx_sc_s* s, const T* pop, T* swap, double* n,
                             double* psi, double* u, bool wf, hipEvent_t mid = nullptr) {
    const unsigned grid = grid_for(s->geo.V);
    launch_density_psi(s, dim3(grid), dim3(256), 0,
                       s->ctx->stream, pop, s->geo, phys(s), n, psi, s->walls);
    if (mid) {
      const hipError_t e = hipEventRecord(mid, s->ctx->stream);
      if (e != hipSuccess) return e;
    }
    if (wf)
      hipLaunchKernelGGL((k_fcs<S, ES, T, SHIFT, COLL, true>), dim3(grid), dim3(256), 0,
                         s->ctx->stream, pop, swap, n, psi, s->geo, phys(s), u, s->walls,
                         s->bforce[0], s->bforce[1], s->bforce[2]);
    else
      hipLaunchKernelGGL((k_fcs<S, ES, T, SHIFT, COLL, false>), dim3(grid), dim3(256), 0,
                         s->ctx->stream, pop, swap, n, psi, s->geo, phys(s), u, s->walls,
                         s->bforce[0], s->bforce[1], s->bforce[2]);
    return hipGetLastError();
  }

  static int run(idmx_sc_s* s, void** pop, void** swap, double* n, double* psi,
                 double* u, int nsteps, int flags) {
    const bool wf = (flags & IDMX_RUN_WRITE_FIELDS) != 0;
    // with walls the step is the two-pass one (density and psi of the fluid
    // sites into n / psi, where psi also holds the solid sites' values)
    const bool body = s->bforce[0] != 0.0 || s->bforce[1] != 0.0 || s->bforce[2] != 0.0;
    // walls (with or without a body force) run the fused step with walls; the
    // step that writes n / psi / u at the end of a run is the two-pass one
    // (from 2^19 sites up, or when asked for: on smaller lattices the
    // two-pass step is as fast; measured 0.94x at 64^3, 1.07x at 1024^2,
    // 1.30x at 256^3 and 1.47x at 4096^2)
    const bool fwalls = kFusedWalls && s->walls && (flags & IDMX_RUN_TWO_PASS) == 0 &&
                        s->desc.psi_kind != IDMX_PSI_PROGRAM &&
                        (s->geo.V >= (int64_t(1) << 19) || (flags & IDMX_RUN_FUSED_WALLS) != 0);
    const bool twop = !fwalls && ((flags & IDMX_RUN_TWO_PASS) != 0 || s->walls || body ||
                                  s->desc.psi_kind == IDMX_PSI_PROGRAM);
    if ((twop || fwalls) && (!n || !psi)) {
      set_error(s->walls || body || s->desc.psi_kind == IDMX_PSI_PROGRAM
                    ? "idmx_sc_run: walls, a body force or a psi program need the n and psi "
                      "buffers (two-pass step)"
                    : "idmx_sc_run: the two-pass step needs n and psi buffers");
      return IDMX_E_ARG;
    }
    if (wf && (!n || !psi || !u)) {
      set_error("idmx_sc_run: IDMX_RUN_WRITE_FIELDS needs n, psi and u buffers");
      return IDMX_E_ARG;
    }
    hipStream_t st = s->ctx->stream;
    // per-step profiling: events 2k and 2k+1 open step k's first and second
    // kernel (the fused step has one: its first interval is empty), event
    // 2 nsteps closes the last step
    hipEvent_t* pev = nullptr;
    if (s->prof_ms) {
      const size_t need = 2 * (size_t)nsteps + 1;
      while (s->prof_ev.size() < need) {
        hipEvent_t e;
        IDMX_HIP(hipEventCreate(&e));
        s->prof_ev.push_back(e);
      }
      pev = s->prof_ev.data();
    }
    // the wall masks, from the flags as they are now (outside the timed region)
    if constexpr (kFusedWalls) if (fwalls) {
      if (!s->wmask_dev) {
        set_error("idmx_sc_run: wall masks not allocated (idmx_sc_set_walls)");
        return IDMX_E_NOMEM;
      }
      hipLaunchKernelGGL((k_wall_mask<S>), dim3(grid_for(s->geo.V)), dim3(256), 0, st, s->walls,
                         s->geo, s->wmask_dev);
      IDMX_HIP(hipGetLastError());
    }
    IDMX_HIP(hipEventRecord(s->ctx->ev0, st));
    // pull-form variants: convert the streamed populations to the pull form,
    // step in it, convert back (nsteps + 2 buffer swaps: same parity as
    // nsteps, so the caller's pointer bookkeeping is unchanged)
    const bool pull = COLL < 3 && !twop && !fwalls && pull_form(s->hist) && nsteps > 0;
    // variant 10, the fused step with walls and the fixed Shan-Chen-forcing /
    // isotropy-filter launch keep the populations in the x-shifted storage
    // between their steps (k_fused_dma<..., XSH>): the first of the nxsh
    // steps reads the streamed storage, the last writes it (the step that
    // writes n/psi/u runs on the streamed storage), so a run of one step
    // never leaves it
    const bool xs = fwalls || (COLL < 3 ? s->hist == 10 : (kFixedDma2 || kFixedDma3));
    const int nxsh = xs && !twop ? nsteps - (wf ? 1 : 0) : 0;
    if constexpr (COLL < 3) if (pull) {
      hipLaunchKernelGGL((k_restream<S, T>), dim3(grid_for(s->geo.V)), dim3(256), 0, st,
                         (const T*)*pop, (T*)*swap, s->geo, 1);
      IDMX_HIP(hipGetLastError());
      std::swap(*pop, *swap);
    }
    for (int k = 0; k < nsteps; k++) {
      const bool last = k == nsteps - 1;
      const T* a = (const T*)*pop;
      T* b = (T*)*swap;
      hipError_t e;
      if (pev) IDMX_HIP(hipEventRecord(pev[2 * k], st));
      hipEvent_t mid = pev ? pev[2 * k + 1] : nullptr;
      if constexpr (COLL >= 3) {
        // the Shan-Chen forcing and the isotropy filter: the fixed fused
        // launch, or the two-pass step (walls, a body force, a psi program,
        // IDMX_RUN_TWO_PASS)
        if (twop) {
          e = two_pass(s, a, b, n, psi, u, wf && last, mid);
        } else {
          if (mid) IDMX_HIP(hipEventRecord(mid, st));
          s->xsh = k < nxsh ? (k > 0 ? 1 : 0) | (k < nxsh - 1 ? 2 : 0) : 0;
          if (wf && last) e = fused_fixed<true>(s, a, b, n, psi, u);
          else e = fused_fixed<false>(s, a, b, nullptr, nullptr, nullptr);
          s->xsh = 0;
        }
      } else {
        if (twop || (fwalls && wf && last)) {
          e = two_pass(s, a, b, n, psi, u, wf && last, mid);
        } else if (fwalls) {
          if (mid) IDMX_HIP(hipEventRecord(mid, st));
          s->xsh = k < nxsh ? (k > 0 ? 1 : 0) | (k < nxsh - 1 ? 2 : 0) : 0;
          e = fused_walls(s, a, b, psi);
          s->xsh = 0;
        } else {
          if (mid) IDMX_HIP(hipEventRecord(mid, st));
          s->xsh = k < nxsh ? (k > 0 ? 1 : 0) | (k < nxsh - 1 ? 2 : 0) : 0;
          if (wf && last) e = fused<true>(s, a, b, n, psi, u);
          else e = fused<false>(s, a, b, nullptr, nullptr, nullptr);
          s->xsh = 0;
        }
      }
      if (e != hipSuccess) return hip_fail(e, "idmx_sc_run launch");
      std::swap(*pop, *swap);
    }
    if (pev && nsteps > 0) IDMX_HIP(hipEventRecord(pev[2 * nsteps], st));
    if constexpr (COLL < 3) if (pull) {
      hipLaunchKernelGGL((k_restream<S, T>), dim3(grid_for(s->geo.V)), dim3(256), 0, st,
                         (const T*)*pop, (T*)*swap, s->geo, -1);
      IDMX_HIP(hipGetLastError());
      std::swap(*pop, *swap);
    }
    IDMX_HIP(hipEventRecord(s->ctx->ev1, st));
    IDMX_HIP(hipEventSynchronize(s->ctx->ev1));
    float ms = 0;
    IDMX_HIP(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
    s->last_ms = ms;
    if (pev) {
      for (int k = 0; k < 2 * nsteps; k++)
        IDMX_HIP(hipEventElapsedTime(&s->prof_ms[k], pev[k], pev[k + 1]));
    }
    return IDMX_OK;
  }

  static int density_psi(idmx_sc_s* s, double* n, double* psi, const void* pop) {
    launch_density_psi(s, dim3(grid_for(s->geo.V)), dim3(256),
                       0, s->ctx->stream, (const T*)pop, s->geo, phys(s), n, psi, s->walls);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int fcs(idmx_sc_s* s, void* swap, const void* pop, const double* n,
                 const double* psi, double* u) {
    const unsigned grid = grid_for(s->geo.V);
    if (u)
      hipLaunchKernelGGL((k_fcs<S, ES, T, SHIFT, COLL, true>), dim3(grid), dim3(256), 0,
                         s->ctx->stream, (const T*)pop, (T*)swap, n, psi, s->geo, phys(s), u,
                         s->walls, s->bforce[0], s->bforce[1], s->bforce[2]);
    else
      hipLaunchKernelGGL((k_fcs<S, ES, T, SHIFT, COLL, false>), dim3(grid), dim3(256), 0,
                         s->ctx->stream, (const T*)pop, (T*)swap, n, psi, s->geo, phys(s),
                         nullptr, s->walls, s->bforce[0], s->bforce[1], s->bforce[2]);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int moments(idmx_sc_s* s, const void* pop, double* n, double* psi, double* u) {
    const unsigned grid = grid_for(s->geo.V);
    launch_density_psi(s, dim3(grid), dim3(256), 0,
                       s->ctx->stream, (const T*)pop, s->geo, phys(s), n, psi, s->walls);
    hipLaunchKernelGGL((k_velocity<S, ES, T, SHIFT>), dim3(grid), dim3(256), 0,
                       s->ctx->stream, (const T*)pop, n, psi, s->geo, s->desc.SC_G, u,
                       s->walls, s->bforce[0], s->bforce[1], s->bforce[2]);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  // K_ComputeMoments (legacy, LBM.py:1667-1698): n and u = sum xi f / n.
  static int density_velocity(idmx_sc_s* s, const void* pop, double* n, double* u) {
    const unsigned grid = grid_for(s->geo.V);
    launch_density_psi(s, dim3(grid), dim3(256), 0,
                       s->ctx->stream, (const T*)pop, s->geo, phys(s), n, nullptr, nullptr);
    // G = 0: no force term; n stands in for the (unused) psi argument
    hipLaunchKernelGGL((k_velocity<S, ES, T, SHIFT>), dim3(grid), dim3(256), 0,
                       s->ctx->stream, (const T*)pop, n, n, s->geo, 0.0, u, nullptr, 0.0, 0.0,
                       0.0);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int init_pop(idmx_sc_s* s, void* pop, const double* n, const double* u) {
    hipLaunchKernelGGL((k_init_pop<S, T, SHIFT>), dim3(grid_for(s->geo.V)), dim3(256), 0,
                       s->ctx->stream, (T*)pop, n, u, s->geo, s->desc.rho0);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int init_pop_tab(idmx_sc_s* s, void* pop, const double* n, const double* u,
                          const double* coeffs) {
    EqTab<NBasis<S::D>::value * S::Q> tab;
    for (int i = 0; i < NBasis<S::D>::value * S::Q; i++) tab.c[i] = coeffs[i];
    hipLaunchKernelGGL((k_init_pop_tab<S, T, SHIFT>), dim3(grid_for(s->geo.V)), dim3(256), 0,
                       s->ctx->stream, (T*)pop, n, u, s->geo, s->desc.rho0, tab);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int mass_mom(idmx_sc_s* s, const void* pop, double* mass, double* mom) {
    const unsigned grid = std::min<unsigned>(grid_for(s->geo.V), 1024);
    idmx_ctx_s* c = s->ctx;
    if ((size_t)grid * 4 > c->partials_cap) {
      set_error("idmx_sc_mass_momentum: reduction scratch too small");
      return IDMX_E_NOMEM;
    }
    hipLaunchKernelGGL((k_mass_mom_partial<S, T, SHIFT>), dim3(grid), dim3(256), 0,
                       c->stream, (const T*)pop, s->geo, s->desc.rho0, c->partials);
    IDMX_HIP(hipGetLastError());
    static thread_local double host[1024 * 4];
    IDMX_HIP(hipMemcpyAsync(host, c->partials, sizeof(double) * 4 * grid,
                            hipMemcpyDeviceToHost, c->stream));
    IDMX_HIP(hipStreamSynchronize(c->stream));
    double acc[4] = {0, 0, 0, 0};
    for (unsigned b = 0; b < grid; b++)
      for (int k = 0; k < 4; k++) acc[k] += host[b * 4 + k];
    *mass = acc[0];
    for (int d = 0; d < S::D; d++) mom[d] = acc[1 + d];
    return IDMX_OK;
  }

  static int to_f64(idmx_sc_s* s, const void* pop, double* dst) {
    hipLaunchKernelGGL((k_pop_to_f64<S, T, SHIFT>), dim3(grid_for(s->geo.V)), dim3(256), 0,
                       s->ctx->stream, (const T*)pop, dst, s->geo, s->desc.rho0);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  static int from_f64(idmx_sc_s* s, void* pop, const double* src) {
    hipLaunchKernelGGL((k_pop_from_f64<S, T, SHIFT>), dim3(grid_for(s->geo.V)), dim3(256), 0,
                       s->ctx->stream, (T*)pop, src, s->geo, s->desc.rho0);
    IDMX_HIP(hipGetLastError());
    return IDMX_OK;
  }

  // Whether launch variant v can run the tile tx x ty (the LDS staging of
  // variants 9 and 10 has to fit one CU).
  static int launch_ok(int tx, int ty, int v) {
    if (v != 9 && v != 10) return 1;
    if constexpr (S::D == 2) {
      switch (tx) {
        case 64: return DmaShape<S, ES, T, 64, 1, COLL>::fits;
        case 128: return DmaShape<S, ES, T, 128, 1, COLL>::fits;
        default: return DmaShape<S, ES, T, 256, 1, COLL>::fits;
      }
    } else {
      switch (tx * 100 + ty) {
        case 6404: return DmaShape<S, ES, T, 64, 4, COLL>::fits;
        case 12802: return DmaShape<S, ES, T, 128, 2, COLL>::fits;
        case 25602: return DmaShape<S, ES, T, 256, 2, COLL>::fits;
        default: return DmaShape<S, ES, T, 32, 8, COLL>::fits;
      }
    }
  }

  static const SimOps* ops() {
    static const SimOps o = {run, density_psi, fcs, moments, density_velocity, init_pop,
                             mass_mom, to_f64, from_f64, launch_ok, init_pop_tab};
    return &o;
  }
};

// Tables over the collision for one (S, ES), FP64 and FP32-shifted. The two
// precisions are instantiated in separate translation units: the FP64 ones
// are compiled without floating-point contraction (no fused multiply-add),
// which is how the reference's CTypes backend computes, so that the FP64
// trajectory stays within 1e-13 of it over 10^4 steps; with contraction the
// two drift apart by 2e-11 over the same run. The FP32-shifted step keeps
// contraction (its accuracy is set by FP32 storage).
// coll: 0 BGK, 1 MRT with dense tables, 2 MRT with a diagonal S, 3 BGK with
// the Shan-Chen forcing, 4 BGK followed by the isotropy filter (D2Q9 and
// FP64 only; 3 and 4 on one fixed fused launch or the two-pass step).
template <class S, class ES>
const SimOps* ops_for_f64(int coll) {
  if (coll == 4) {
    if constexpr (S::Q == 9) return Impl<S, ES, double, false, 4>::ops();
    else return nullptr;
  }
  if (coll == 0) return Impl<S, ES, double, false, 0>::ops();
  if (coll == 1) return Impl<S, ES, double, false, 1>::ops();
  if (coll == 3) return Impl<S, ES, double, false, 3>::ops();
  return Impl<S, ES, double, false, 2>::ops();
}
template <class S, class ES>
const SimOps* ops_for_f32(int coll) {
  if (coll == 4) return nullptr;
  if (coll == 0) return Impl<S, ES, float, true, 0>::ops();
  if (coll == 1) return Impl<S, ES, float, true, 1>::ops();
  if (coll == 3) return Impl<S, ES, float, true, 3>::ops();
  return Impl<S, ES, float, true, 2>::ops();
}

}  // namespace idmx
