// batch_kernels.h -- one launch advances many independent D2Q9 Shan-Chen
// domains of different sizes and couplings (a (G, R, L) parameter sweep).
//
// The reference bakes V, L_i and STR_i into every kernel as #defines
// (idpy/LBM/LBM.py:542-569, idpy/IdpyCode/IdpyCode.py:770-781), so each domain
// size is a new compile and a new launch. Here the geometry is data: a table
// of domain descriptors and a list of work items (domain, x0, y0, rows).
// Persistent blocks stride over the item list. Two item kinds: a march of
// the fused step over H rows of a TX-wide chunk (the single-domain kernel's
// algorithm: LDS psi ring with halo, populations kept in registers, push
// streaming; large domains), or one BX x BY tile with its halo (small
// domains, where a march would be a chain of dependent row loads). The
// number of live items is read from device memory, so the host can drop
// converged domains between graph replays without re-capturing the graph.
#pragma once
#include "sc_kernels.h"

namespace idmx {

struct BatchDom {
  int L0, L1;
  int64_t pop_off;   // element offset of the domain's Q*V populations
  int64_t site_off;  // offset of the domain's V sites in per-site arrays
  double G, omega, rho0;
  int psi_kind;
  int pad;
};

struct BatchItem {
  int dom, x0, m0, Hs;
};

// One TX-wide, Hs-row march of the 2D fused step on one domain (R = stencil
// reach). Same structure as k_fused's 2D path; see there for the ring and
// register-history logic.
template <class ES, class T, bool SHIFT, int TX, bool PULL = false>
__device__ inline void march2d(const T* __restrict__ pop, T* __restrict__ pop_swap,
                               const Geom& geo, const Phys<T>& P, int x0, int m0,
                               int Hs, T (*lds)[TX + 2 * ES::R]) {
  using S = D2Q9;
  constexpr int Q = S::Q, R = ES::R;
  constexpr int NS = 2 * R + 2;
  const int Lm = geo.L1;
  const int64_t PS = geo.L0;
  const int tid = threadIdx.x;
  const int x = x0 + tid;
  const bool own_valid = x < geo.L0;
  const int64_t own_off = x % geo.L0;
  const bool has_halo = tid < 2 * R;
  const int hx = tid < R ? tid : TX + tid;
  const int64_t halo_off = has_halo ? ((x0 - R + hx) % geo.L0 + geo.L0) % geo.L0 : 0;

  T fh[R + 1][Q];
  T dnh[R + 1];
  for (int j = 0; j < Hs + 2 * R; j++) {
    const int mj = ((m0 - R + j) % Lm + Lm) % Lm;
    const int64_t base = (int64_t)mj * PS;
    const int slot = j % NS;
    {
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++)
        fh[R][q] = pop[(size_t)q * geo.V +
                       (PULL ? pull_src<S>(geo, q, (int)own_off, 0, mj) : base + own_off)];
#pragma unroll
      for (int q = 0; q < Q; q++) s += fh[R][q];
      dnh[R] = s;
      lds[slot][R + tid] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
    }
    if (has_halo) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++)
        s += pop[(size_t)q * geo.V +
                 (PULL ? pull_src<S>(geo, q, (int)halo_off, 0, mj) : base + halo_off)];
      lds[slot][hx] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
    }
    __syncthreads();
    if (j >= 2 * R && own_valid) {
      int sl[2 * R + 1];
#pragma unroll
      for (int k = 0; k <= 2 * R; k++) sl[k] = (j - 2 * R + k) % NS;
      const int cx = R + tid;
      T gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int e = 0; e < ES::QE; e++) {
        const T v = lds[sl[R + ES::ey[e]]][cx + ES::ex[e]];
        const T w = T(ES::w[e]);
        if (ES::ex[e] != 0) gx += T(ES::ex[e]) * w * v;
        if (ES::ey[e] != 0) gy += T(ES::ey[e]) * w * v;
      }
      const T lpsi = lds[sl[R]][cx];
      T fo[Q], u[3];
      collide<S, T, SHIFT, 0>(fh[0], dnh[0], gx, gy, gz, lpsi, P, fo, u);
      if constexpr (PULL) local_store<S, T>(pop_swap, geo, x, m0 + j - 2 * R, 0, fo);
      else push_store<S, T>(pop_swap, geo, x, m0 + j - 2 * R, 0, fo);
    }
#pragma unroll
    for (int k = 0; k < R; k++) {
      dnh[k] = dnh[k + 1];
#pragma unroll
      for (int q = 0; q < Q; q++) fh[k][q] = fh[k + 1][q];
    }
  }
}

template <class T>
__device__ inline Phys<T> dom_phys(const BatchDom& d) {
  Phys<T> P;
  P.G = (T)d.G;
  P.omega = (T)d.omega;
  P.rho0 = (T)d.rho0;
  P.psi_kind = d.psi_kind;
  P.mrtA = P.mrtB = P.mrtG = nullptr;
  return P;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8),
// each with its own L2. This permutation of [0, G) hands XCD x the
// contiguous virtual indices [x q + min(x, r), ...), q + (x < r) of them
// (q = G / 8, r = G % 8), so that neighbouring work items, which re-read
// each other's edge populations, mostly run on one XCD. Any grid size works.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, G = gridDim.x;
  const int x = b & 7, q = G >> 3, r = G & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <class ES, class T, bool SHIFT, int TX, bool PULL = false>
__global__ void __launch_bounds__(TX)
k_batched(const T* __restrict__ pop_all, T* __restrict__ swap_all,
          const BatchDom* __restrict__ doms, const BatchItem* __restrict__ items,
          const int* __restrict__ n_items) {
  __shared__ T lds[2 * ES::R + 2][TX + 2 * ES::R];
  const int ni = *n_items;
  for (int it = xcd_block(); it < ni; it += gridDim.x) {
    const BatchItem w = items[it];
    const BatchDom d = doms[w.dom];
    const Geom geo{d.L0, d.L1, 1, (int64_t)d.L0 * d.L1};
    const Phys<T> P = dom_phys<T>(d);
    march2d<ES, T, SHIFT, TX, PULL>(pop_all + d.pop_off, swap_all + d.pop_off, geo, P, w.x0,
                                    w.m0, w.Hs, lds);
    __syncthreads();
  }
}

// Non-negative modulo for coordinates any number of periods out of range
// (a tile can be wider than a small domain).
__device__ inline int pmod(int v, int L) {
  const int r = v % L;
  return r < 0 ? r + L : r;
}

// One BX x BY tile of one domain per work item, one site per thread. The
// tile's psi and an R-wide ring of halo psi are stored in LDS (the halo
// densities are summed from their populations, which neighbouring tiles also
// read, so those loads mostly hit L2), one barrier, then every site of the
// tile collides and push-streams. Every load of an item is independent of
// the others, so a step costs about two memory latencies however few rows a
// domain has; the marching kernel above pays H + 2R dependent row loads per
// item and suits large domains instead.
//
// A thread whose site lies past the domain edge (a tile wider or taller than
// what is left of the domain) still loads the periodic image of its
// position: that is the halo its valid neighbours need. Only sites inside
// the domain collide and store.
template <class ES, class T, bool SHIFT, int BX, int BY, bool PULL = false>
__global__ void __launch_bounds__(BX * BY)
k_batched_tile(const T* __restrict__ pop_all, T* __restrict__ swap_all,
               const BatchDom* __restrict__ doms, const BatchItem* __restrict__ items,
               const int* __restrict__ n_items) {
  using S = D2Q9;
  constexpr int Q = S::Q, R = ES::R;
  constexpr int WX = BX + 2 * R, WY = BY + 2 * R, NT = BX * BY;
  constexpr int NH = WX * WY - NT;  // halo ring: 2R rows of WX, 2R columns of BY
  __shared__ T lds[WY][WX];
  const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
  const int ni = *n_items;
  for (int it = xcd_block(); it < ni; it += gridDim.x) {
    const BatchItem w = items[it];
    const BatchDom d = doms[w.dom];
    const Geom geo{d.L0, d.L1, 1, (int64_t)d.L0 * d.L1};
    const Phys<T> P = dom_phys<T>(d);
    const T* pop = pop_all + d.pop_off;
    const int x = w.x0 + tx, y = w.m0 + ty;
    T f[Q];
    T s = 0;
    {
      const int xs = pmod(x, d.L0), ys = pmod(y, d.L1);
      const int64_t g = (int64_t)ys * d.L0 + xs;
#pragma unroll
      for (int q = 0; q < Q; q++)
        f[q] = pop[(size_t)q * geo.V + (PULL ? pull_src<S>(geo, q, xs, 0, ys) : g)];
#pragma unroll
      for (int q = 0; q < Q; q++) s += f[q];
      lds[ty + R][tx + R] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
    }
    for (int h = threadIdx.x; h < NH; h += NT) {
      int hx, hy;
      if (h < 2 * R * WX) {  // R rows above the tile, then R rows below
        const int r = h / WX;
        hx = h - r * WX;
        hy = r < R ? r : BY + r;
      } else {  // R columns left and right of each tile row
        const int k = h - 2 * R * WX;
        const int r = k / (2 * R), c = k - r * (2 * R);
        hy = R + r;
        hx = c < R ? c : BX + c;
      }
      const int hxs = pmod(w.x0 - R + hx, d.L0), hys = pmod(w.m0 - R + hy, d.L1);
      const int64_t g = (int64_t)hys * d.L0 + hxs;
      T sh = 0;
#pragma unroll
      for (int q = 0; q < Q; q++)
        sh += pop[(size_t)q * geo.V + (PULL ? pull_src<S>(geo, q, hxs, 0, hys) : g)];
      lds[hy][hx] = psi_of(P.psi_kind, SHIFT ? P.rho0 + sh : sh);
    }
    __syncthreads();
    if (x < d.L0 && y < d.L1) {
      T gx = 0, gy = 0;
#pragma unroll
      for (int e = 0; e < ES::QE; e++) {
        const T v = lds[ty + R + ES::ey[e]][tx + R + ES::ex[e]];
        const T wgt = T(ES::w[e]);
        if (ES::ex[e] != 0) gx += T(ES::ex[e]) * wgt * v;
        if (ES::ey[e] != 0) gy += T(ES::ey[e]) * wgt * v;
      }
      T fo[Q], u[3];
      collide<S, T, SHIFT, 0>(f, s, gx, gy, T(0), lds[ty + R][tx + R], P, fo, u);
      if constexpr (PULL) local_store<S, T>(swap_all + d.pop_off, geo, x, y, 0, fo);
      else push_store<S, T>(swap_all + d.pop_off, geo, x, y, 0, fo);
    }
    __syncthreads();
  }
}

// Two steps per pass of one BX x BY tile (E4 force, reach 1; push form).
// The tile's second-step collisions need the first step's post-collision
// populations on the tile plus a 2-site ring (C0), which need the densities
// on the tile plus a 3-site ring (L0); the first step's results never leave
// LDS, so HBM sees each population read once and written once per two
// steps (the ring reads come mostly from L2). Every site is updated by the
// same arithmetic, in the same order, as two passes of k_batched_tile, so
// the two agree bit for bit.
//   phase A: load f(0) of L0 (each thread keeps the values of its L0 sites
//            in registers), n(0), psi(0) -> LDS;
//   phase B: collide the C0 sites (force from psi(0)) -> f*(0) in LDS;
//   phase C: f(1)(y) = f*(0)(y - c_q) on the tile plus a 1-site ring, n(1),
//            psi(1) -> LDS;
//   phase D: collide the tile's own sites and push-store f*(1).
// xsh (kind 5): bit 0, phase A reads the x-shifted storage (population q of
// site x kept at x - c_qx); bit 1, phase D writes it, so that the stores of
// every population cover the tile's own x range. idmx_batch_run's graph
// enters and leaves it within each replay.
template <int BX, int BY>
struct Tile2Shape {
  static constexpr int NT = BX * BY;
  static constexpr int LW = BX + 6, LH = BY + 6, NL = LW * LH;  // L0
  static constexpr int CW = BX + 4, CH = BY + 4, NC = CW * CH;  // C0
  static constexpr int PW = BX + 2, PH = BY + 2, NP = PW * PH;  // tile + 1 ring
  static constexpr int K = (NL + NT - 1) / NT;                  // L0 sites per thread
};

// Phase A's loads: f(0) of the item's L0 sites into registers (each thread
// keeps the values of its L0 sites).
template <class T, int BX, int BY>
__device__ __forceinline__ void tile2_load(const T* __restrict__ pop, const BatchItem& w,
                                           const BatchDom& d,
                                           T (&f0)[Tile2Shape<BX, BY>::K][D2Q9::Q], bool rsh) {
  using S = D2Q9;
  using SH = Tile2Shape<BX, BY>;
  const int64_t V = (int64_t)d.L0 * d.L1;
#pragma unroll
  for (int k = 0; k < SH::K; k++) {
    const int sidx = threadIdx.x + k * SH::NT;
    if (sidx < SH::NL) {
      const int lx = sidx % SH::LW, ly = sidx / SH::LW;
      const int xs = pmod(w.x0 - 3 + lx, d.L0), ys = pmod(w.m0 - 3 + ly, d.L1);
      const int64_t g = (int64_t)ys * d.L0 + xs;
      // the x-shifted storage keeps population q of site x at x - c_qx
      const int64_t gp = rsh ? (int64_t)ys * d.L0 + pmod(xs - 1, d.L0) : g;
      const int64_t gm = rsh ? (int64_t)ys * d.L0 + pmod(xs + 1, d.L0) : g;
#pragma unroll
      for (int q = 0; q < S::Q; q++)
        f0[k][q] = pop[(size_t)q * V + (S::cx[q] > 0 ? gp : S::cx[q] < 0 ? gm : g)];
    }
  }
}

// Phases A (from the loaded values) to D of one item. The barriers wait for
// LDS only: no global memory written in a launch is read in it, and the
// push stores stay in flight through the next barrier.
template <class T, bool SHIFT, int BX, int BY>
__device__ __forceinline__ void tile2_body(
    const T (&f0)[Tile2Shape<BX, BY>::K][D2Q9::Q], const BatchItem& w, const BatchDom& d,
    T* __restrict__ swap_all, T (*a0)[Tile2Shape<BX, BY>::NC],
    T (*psi0)[Tile2Shape<BX, BY>::LW], T (*psi1)[Tile2Shape<BX, BY>::PW], bool wsh) {
  using S = D2Q9;
  using ES = D2E4;
  using SH = Tile2Shape<BX, BY>;
  constexpr int Q = S::Q, NT = SH::NT, K = SH::K;
  static_assert(ES::R == 1, "two steps per pass: reach-1 force stencil");
  const int tid = threadIdx.x;
  const int tx = tid % BX, ty = tid / BX;
  const Geom geo{d.L0, d.L1, 1, (int64_t)d.L0 * d.L1};
  const Phys<T> P = dom_phys<T>(d);
  // phase A: n(0), psi(0) of L0 = [x0 - 3, x0 + BX + 3) x [y0 - 3, y0 + BY + 3)
  T s0[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int sidx = tid + k * NT;
    if (sidx < SH::NL) {
      const int lx = sidx % SH::LW, ly = sidx / SH::LW;
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++) s += f0[k][q];
      s0[k] = s;
      psi0[ly][lx] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
    }
  }
  lds_barrier();
  // phase B: the C0 sites among this thread's L0 sites
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int sidx = tid + k * NT;
    const int lx = sidx % SH::LW, ly = sidx / SH::LW;
    if (sidx < SH::NL && lx >= 1 && lx < SH::LW - 1 && ly >= 1 && ly < SH::LH - 1) {
      T gx = 0, gy = 0;
#pragma unroll
      for (int e = 0; e < ES::QE; e++) {
        const T v = psi0[ly + ES::ey[e]][lx + ES::ex[e]];
        const T wgt = T(ES::w[e]);
        if (ES::ex[e] != 0) gx += T(ES::ex[e]) * wgt * v;
        if (ES::ey[e] != 0) gy += T(ES::ey[e]) * wgt * v;
      }
      T fo[Q], u[3];
      collide<S, T, SHIFT, 0>(f0[k], s0[k], gx, gy, T(0), psi0[ly][lx], P, fo, u);
      const int c = (ly - 1) * SH::CW + (lx - 1);
#pragma unroll
      for (int q = 0; q < Q; q++) a0[q][c] = fo[q];
    }
  }
  lds_barrier();
  // phase C: n(1), psi(1) on the tile plus a 1-site ring (P region)
  for (int sidx = tid; sidx < SH::NP; sidx += NT) {
    const int px = sidx % SH::PW, py = sidx / SH::PW;
    T s = 0;
#pragma unroll
    for (int q = 0; q < Q; q++)
      s += a0[q][(py + 1 - S::cy[q]) * SH::CW + (px + 1 - S::cx[q])];
    psi1[py][px] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
  }
  lds_barrier();
  // phase D: the tile's own sites
  const int x = w.x0 + tx, y = w.m0 + ty;
  if (x < d.L0 && y < d.L1) {
    T f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = a0[q][(ty + 2 - S::cy[q]) * SH::CW + (tx + 2 - S::cx[q])];
    T s = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) s += f[q];
    T gx = 0, gy = 0;
#pragma unroll
    for (int e = 0; e < ES::QE; e++) {
      const T v = psi1[ty + 1 + ES::ey[e]][tx + 1 + ES::ex[e]];
      const T wgt = T(ES::w[e]);
      if (ES::ex[e] != 0) gx += T(ES::ex[e]) * wgt * v;
      if (ES::ey[e] != 0) gy += T(ES::ey[e]) * wgt * v;
    }
    T fo[Q], u[3];
    collide<S, T, SHIFT, 0>(f, s, gx, gy, T(0), psi1[ty + 1][tx + 1], P, fo, u);
    if (wsh) {
      // x-shifted storage: f*_q of site x goes to x + c_q, kept at its x
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int yt = S::cy[q] ? wrap(y + S::cy[q], geo.L1) : y;
        pop_st_nt(swap_all + d.pop_off, q, geo.V, (int64_t)yt * d.L0 + x, fo[q]);
      }
    } else {
      push_store<S, T>(swap_all + d.pop_off, geo, x, y, 0, fo);
    }
  }
  lds_barrier();
}

template <class T, bool SHIFT, int BX, int BY>
__global__ void __launch_bounds__(BX * BY)
k_batched_tile2(const T* __restrict__ pop_all, T* __restrict__ swap_all,
                const BatchDom* __restrict__ doms, const BatchItem* __restrict__ items,
                const int* __restrict__ n_items, int xsh) {
  using SH = Tile2Shape<BX, BY>;
  __shared__ T a0[D2Q9::Q][SH::NC];      // f*(0) on C0
  __shared__ T psi0[SH::LH][SH::LW];
  __shared__ T psi1[SH::PH][SH::PW];
  const int ni = *n_items;
  for (int it = xcd_block(); it < ni; it += gridDim.x) {
    const BatchItem w = items[it];
    const BatchDom d = doms[w.dom];
    T f0[SH::K][D2Q9::Q];
    tile2_load<T, BX, BY>(pop_all + d.pop_off, w, d, f0, (xsh & 1) != 0);
    tile2_body<T, SHIFT, BX, BY>(f0, w, d, swap_all, a0, psi0, psi1, (xsh & 2) != 0);
  }
}

// Per-domain max |n - n_prev| over the live items (FP64), n_prev <- n. An
// item covers W columns (its x-chunk or tile width) and Hs rows; 256
// threads stride over its sites. dmax holds the bit patterns of non-negative
// doubles, whose unsigned order is their numeric order, so a 64-bit atomic
// max reduces them.
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_batched_dn(const T* __restrict__ pop_all, const BatchDom* __restrict__ doms,
             const BatchItem* __restrict__ items, const int* __restrict__ n_items, int W,
             double* __restrict__ n_prev, unsigned long long* __restrict__ dmax) {
  const int ni = *n_items;
  __shared__ double part[4];
  for (int it = blockIdx.x; it < ni; it += gridDim.x) {
    const BatchItem w = items[it];
    const BatchDom d = doms[w.dom];
    const int64_t V = (int64_t)d.L0 * d.L1;
    const T* pop = pop_all + d.pop_off;
    double m = 0.0;
    for (int k = threadIdx.x; k < W * w.Hs; k += 256) {
      const int r = k / W, x = w.x0 + (k - r * W);
      if (x >= d.L0 || w.m0 + r >= d.L1) continue;
      const int64_t g = (int64_t)(w.m0 + r) * d.L0 + x;
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 9; q++) s += (double)pop[(size_t)q * V + g];
      const double n = SHIFT ? s + d.rho0 : s;
      double* np = n_prev + d.site_off + g;
      const double dd = fabs(n - *np);
      m = dd == dd ? fmax(m, dd) : INFINITY;  // a NaN state reports +inf
      *np = n;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double r = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
      atomicMax(dmax + w.dom, (unsigned long long)__double_as_longlong(r));
    }
    __syncthreads();
  }
}

// f = f_eq(n, 0) (shifted when SHIFT) for one domain, n in FP64.
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_batched_init(T* __restrict__ pop, const double* __restrict__ n, int64_t V, double rho0) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double f = D2Q9::w[q] * n[g];
      if (SHIFT) f -= D2Q9::w[q] * rho0;
      pop[(size_t)q * V + g] = (T)f;
    }
  }
}

// n = sum f (+ rho0) in FP64 for one domain.
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_batched_density(const T* __restrict__ pop, double* __restrict__ n, int64_t V, double rho0) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) s += (double)pop[(size_t)q * V + g];
    n[g] = SHIFT ? s + rho0 : s;
  }
}

// Population copy between two batches of the same domains (e.g. FP32
// shifted -> FP64), unshifting / reshifting on the way.
template <class TI, bool SI, class TO, bool SO>
__global__ void __launch_bounds__(256)
k_batched_convert(const TI* __restrict__ src, TO* __restrict__ dst, int64_t V, double rho_in,
                  double rho_out) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double f = (double)src[(size_t)q * V + g];
      if (SI) f += D2Q9::w[q] * rho_in;
      if (SO) f -= D2Q9::w[q] * rho_out;
      dst[(size_t)q * V + g] = (TO)f;
    }
  }
}

// Rest equilibrium f = w n of the source batch's density (summed in FP64),
// written in the destination's storage: drops the momentum the source
// carried. FP32 rounding gives a droplet a small net momentum, which FP64
// then conserves exactly, so the droplet would translate for ever.
template <class TI, bool SI, class TO, bool SO>
__global__ void __launch_bounds__(256)
k_batched_rest(const TI* __restrict__ src, TO* __restrict__ dst, int64_t V, double rho_in,
               double rho_out, double* __restrict__ n_ref) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) s += (double)src[(size_t)q * V + g];
    const double n = SI ? s + rho_in : s;
#pragma unroll
    for (int q = 0; q < 9; q++)
      dst[(size_t)q * V + g] = (TO)(D2Q9::w[q] * (SO ? n - rho_out : n));
    n_ref[g] = n;
  }
}

// Projection of one domain's density onto the fields symmetric about its
// centre c = (L0/2, L1/2) (integer division): the average over the mirror
// images x -> 2 c0 - x, y -> 2 c1 - y (mod L) and, on square domains, the
// transpose, i.e. the dihedral group of the square (8 images) or of the
// rectangle (4), plus a uniform shift (0 keeps the mass: averaging over a
// permutation group of the sites does). The populations become the rest
// equilibrium of the result. n holds the current density (read only); the
// result goes to pop and to the check reference n_ref.
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_batched_symmetrize(const double* __restrict__ n, T* __restrict__ pop, int L0, int L1,
                     double rho0, double shift, double* __restrict__ n_ref) {
  const int64_t V = (int64_t)L0 * L1;
  const int c0 = L0 / 2, c1 = L1 / 2;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(g / L0), x = (int)(g - (int64_t)y * L0);
    const int xr = pmod(2 * c0 - x, L0), yr = pmod(2 * c1 - y, L1);
    double s = n[(int64_t)y * L0 + x] + n[(int64_t)y * L0 + xr] + n[(int64_t)yr * L0 + x] +
               n[(int64_t)yr * L0 + xr];
    double m = 4.0;
    if (L0 == L1) {
      s += n[(int64_t)x * L0 + y] + n[(int64_t)x * L0 + yr] + n[(int64_t)xr * L0 + y] +
           n[(int64_t)xr * L0 + yr];
      m = 8.0;
    }
    const double v = s / m + shift;
#pragma unroll
    for (int q = 0; q < 9; q++)
      pop[(size_t)q * V + g] = (T)(D2Q9::w[q] * (SHIFT ? v - rho0 : v));
    n_ref[g] = v;
  }
}

// Populations of one domain at rest under the Guo forcing, from its density n
// (FP64, periodic). The collision takes the velocity u = (sum_q xi_q f_q +
// F/2) / n, so f = w n (zero first moment) puts fluid at rest only where F
// vanishes: at an interface it starts the fluid at F / (2n), a kick the
// slowly damped modes of the domain then carry for ~1e4-1e5 steps. Here
// f = f_eq(n, v) with v = -F / (2n), i.e. first moment -F/2, so the
// collision sees u = 0 everywhere. F = -G psi(x) sum_k W_k e_k psi(x + e_k)
// is the domain's Shan-Chen force from n. The result goes to pop (shifted
// by w rho0 when SHIFT).
template <class ES, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_batched_force_rest(const double* __restrict__ n, T* __restrict__ pop, int L0, int L1,
                     double G, int psi_kind, double rho0) {
  using S = D2Q9;
  const int64_t V = (int64_t)L0 * L1;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(g / L0), x = (int)(g - (int64_t)y * L0);
    const double nn = n[g];
    double gx = 0.0, gy = 0.0;
#pragma unroll
    for (int k = 0; k < ES::QE; k++) {
      const int xs = pmod(x + ES::ex[k], L0), ys = pmod(y + ES::ey[k], L1);
      const double p = psi_of(psi_kind, n[(int64_t)ys * L0 + xs]);
      gx += ES::w[k] * ES::ex[k] * p;
      gy += ES::w[k] * ES::ey[k] * p;
    }
    const double mg = -G * psi_of(psi_kind, nn);
    const double vx = -0.5 * mg * gx / nn, vy = -0.5 * mg * gy / nn;
    const double vv = vx * vx + vy * vy;
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const double xv = S::cx[q] * vx + S::cy[q] * vy;
      double f = S::w[q] * nn * (1.0 + kCM2 * xv + 0.5 * kCM4 * xv * xv - 0.5 * kCM2 * vv);
      if (SHIFT) f -= S::w[q] * rho0;
      pop[(size_t)q * V + g] = (T)f;
    }
  }
}

// Conversion of every live domain between streamed populations and the pull
// form (see k_restream): dir = +1 to the pull form, -1 back. Items tile the
// live domains, W columns by Hs rows each, 256 threads per item.
template <class T>
__global__ void __launch_bounds__(256)
k_batched_restream(const T* __restrict__ src_all, T* __restrict__ dst_all,
                   const BatchDom* __restrict__ doms, const BatchItem* __restrict__ items,
                   const int* __restrict__ n_items, int W, int dir) {
  const int ni = *n_items;
  for (int it = blockIdx.x; it < ni; it += gridDim.x) {
    const BatchItem w = items[it];
    const BatchDom d = doms[w.dom];
    const Geom geo{d.L0, d.L1, 1, (int64_t)d.L0 * d.L1};
    const T* src = src_all + d.pop_off;
    T* dst = dst_all + d.pop_off;
    for (int k = threadIdx.x; k < W * w.Hs; k += 256) {
      const int r = k / W, x = w.x0 + (k - r * W), y = w.m0 + r;
      if (x >= d.L0 || y >= d.L1) continue;
      const int64_t g = (int64_t)y * d.L0 + x;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const int xs = D2Q9::cx[q] ? wrap(x + dir * D2Q9::cx[q], d.L0) : x;
        const int ys = D2Q9::cy[q] ? wrap(y + dir * D2Q9::cy[q], d.L1) : y;
        dst[(size_t)q * geo.V + g] = src[(size_t)q * geo.V + (int64_t)ys * d.L0 + xs];
      }
    }
  }
}

// Per-block partial sums of n over V sites (FP64).
static __global__ void __launch_bounds__(256)
k_batched_sum_partial(const double* __restrict__ n, int64_t V, double* __restrict__ out) {
  double s = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x)
    s += n[g];
  s = block_reduce<0>(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace idmx
