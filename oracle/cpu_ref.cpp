// cpu_ref.cpp -- CPU restatement of idpy's Shan-Chen multiphase LBM step.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle and the CPU
// baseline; it is never linked into, loaded by, or called from the product
// library (libidmx.so). Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.
//
// It restates, as a serial C++ loop over lattice sites (the CTypes backend runs
// every kernel as `for(g_tid=0; g_tid<N_THREADS; g_tid++)`,
// idpy/IdpyCode/IdpyCode.py:643-664), the arithmetic of:
//   * the legacy class idpy.LBM.LBM.ShanChenMultiPhase (idpy/LBM/LBM.py:806-894)
//     - K_ComputeMoments          idpy/LBM/LBM.py:1667-1698
//     - K_ComputePsi              idpy/LBM/LBM.py:1613-1631
//     - K_Collision_ShanChenGuoMultiPhase idpy/LBM/LBM.py:1529-1611
//     - K_StreamPeriodic (pull)   idpy/LBM/LBM.py:1633-1665
//     - M_SwapPop                 idpy/LBM/LBM.py:1921-1933
//   * the current class idpy.LBM.MultiPhase.ShanChenMultiPhase.MainLoop
//     (idpy/LBM/MultiPhase.py:1015-1097)
//     - K_ComputeDensityPsiMeta   idpy/LBM/MultiPhaseKernelsMeta.py:107-167
//     - K_ForceCollideStreamSCMPMeta (push) idpy/LBM/MultiPhaseKernelsMeta.py:169-345
//   * MainLoopMRT (idpy/LBM/MultiPhase.py:1311-1402) with
//     K_ForceCollideStreamSCMP_MRTMeta idpy/LBM/MultiPhaseKernelsMeta.py:891-1075
//     f* = M^-1 [S M f_eq + (1-S) M f] + M'^-1 (1-S/2) M' F_guo
//     (idpy/LBM/Collisions.py:96-160, idpy/LBM/Forcings.py:129-181); the three
//     Q x Q products are formed once on the host by oracle/hermite_mrt.py.
//   * MainLoopWalls (idpy/LBM/MultiPhase.py:1226-1309): the same step on the
//     fluid sites only (walls[g] == 1), K_ComputeDensityPsiWallsMeta
//     (MultiPhaseKernelsMeta.py:42-104) and K_ForceCollideStreamWallsSCMPMeta
//     (:524-716), whose push stream reflects a population aimed at a solid
//     site (walls == 0) back into the opposite direction at its source site
//     (SRTCollisionPlusGuoPushStreamWallsCode, idpy/LBM/Collisions.py:835-850);
//   * MainLoopSimpleWalls / MainLoopSimpleWallsGravity (idpy/LBM/MultiPhase.py:
//     351-465, 888-1010): the legacy kernels on the flagged sites, pull
//     streaming, then K_HalfWayBounceBack (idpy/LBM/LBMKernels.py:151-184)
//     and a body force F[1] += g; MainLoopSimpleSC / MainLoopSimpleWallsSC
//     (MultiPhase.py:257-349, 467-581) with the Shan-Chen forcing collision
//     (K_Collision_ShanChenMultiPhase[Walls], MultiPhaseKernels.py:132-304);
//   * the same MainLoop with the current class's printed 15-digit
//     coefficients (evalf() / str() of the sympy expansion,
//     idpy/LBM/Collisions.py:251-273), evaluated as the printed sums
//     (Terms, cpuref_meta_run_terms / cpuref_init_pop_terms);
//   * MainLoopIsoFilter (idpy/LBM/MultiPhase.py:1099-1223): the D2Q9 step
//     with K_IsotropyFilter (idpy/LBM/LBMKernelsMeta.py:895-1080) between
//     the collision and the push stream;
//   * ComputeMomentsMeta (idpy/LBM/MultiPhase.py:1565-1662,
//     K_ComputeVelocityAfterForceSCMPMeta MultiPhaseKernelsMeta.py:1077-1219)
//   * InitPopulations (K_InitPopulations idpy/LBM/LBM.py:1700-1743)
//   * InitRadialInterface (legacy, idpy/LBM/LBM.py:1786-1826), InitFlatInterface
//     (:1745-1784), InitCylinderInterface (:1828-1874) and the Meta class's
//     InitSingleFlatInterface (idpy/LBM/MultiPhaseKernels.py:650-684)
//
// Constants are full doubles ("exact" set): the legacy class reads W_list and
// EW_list from memory as doubles (idpy/LBM/LBM.py:1596-1607), and the stencil
// tables of idpy/LBM/LBM.py:74-185 are Python floats.
//
// Compile: g++ -std=c++17 -O3 -ffp-contract=off -fPIC -shared (oracle/Makefile).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- stencils
// Velocity sets: idpy/LBM/LBM.py:74-92.
struct SD2Q9 {
  static constexpr int D = 2, Q = 9;
  static constexpr int xi[9][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},
                                   {-1, 0, 0}, {0, -1, 0}, {1, 1, 0},
                                   {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static constexpr double w[9] = {4. / 9,  1. / 9,  1. / 9,  1. / 9, 1. / 9,
                                  1. / 36, 1. / 36, 1. / 36, 1. / 36};
};
struct SD3Q19 {
  static constexpr int D = 3, Q = 19;
  static constexpr int xi[19][3] = {
      {0, 0, 0},  {1, 0, 0},   {0, 1, 0},  {-1, 0, 0}, {0, -1, 0},
      {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, 1, 0}, {-1, -1, 0},
      {1, -1, 0}, {1, 0, 1},   {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1},
      {0, 1, 1},  {0, -1, 1},  {0, -1, -1}, {0, 1, -1}};
  static constexpr double w[19] = {
      1. / 3,  1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18,
      1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36,
      1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
};
constexpr double kC2 = 1. / 3;            // 'c2' of both velocity sets
constexpr double kCM2 = 1. / kC2;         // CM2 = 1/c2   (LBM.py:542-569)
constexpr double kCM4 = 1. / (kC2 * kC2); // CM4 = 1/c2^2

// Pseudopotential forms accepted through psi_kind (psi_code is a C
// expression in the reference, MultiPhase.py:2120, LBM.py:1244).
// kind 3: a psi_code expression compiled by gcc into a function of n, as the
// CTypes backend compiles it into K_ComputeDensityPsiMeta
// (MultiPhaseKernelsMeta.py:107-167); set with cpuref_set_psi_fn.
double (*g_psi_fn)(double) = nullptr;

inline double psi_of(int kind, double n) {
  switch (kind) {
  case 0: return 1. - exp(-n);   // psi = 1 - exp(-n)
  case 1: return exp(-1. / n);   // psi = exp(-1/n)
  case 3: return g_psi_fn(n);    // psi = psi_code(n)
  default: return n;             // psi = n
  }
}

// Lattice geometry with lexicographic index g = x + y*L0 + z*L0*L1
// (idpy/IdpyCode/IdpyUnroll.py:325-386, idpy/LBM/LBM.py:410-425).
struct Geo {
  int D;
  int L[3];
  int64_t V;
  int64_t idx(int x, int y, int z) const {
    return (int64_t)x + (int64_t)L[0] * ((int64_t)y + (int64_t)L[1] * z);
  }
  void pos(int64_t g, int* p) const {
    p[0] = (int)(g % L[0]);
    p[1] = D > 1 ? (int)((g / L[0]) % L[1]) : 0;
    p[2] = D > 2 ? (int)(g / ((int64_t)L[0] * L[1])) : 0;
  }
  // periodic neighbour, (p + e + L) % L per axis (LBM.py:1573-1575)
  int64_t nb(const int* p, const int* e) const {
    int q[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) q[d] = ((p[d] + e[d]) % L[d] + L[d]) % L[d];
    return idx(q[0], q[1], q[2]);
  }
};

// Population ordering: 'gpu' = SoA pop[g + q*V], 'cpu' = AoS pop[q + g*Q]
// (idpy/LBM/LBM.py:529-537).
struct Ord {
  bool aos;
  int64_t V;
  int Q;
  int64_t operator()(int64_t g, int q) const {
    return aos ? (int64_t)q + g * Q : g + (int64_t)q * V;
  }
};

// E-stencil passed at run time (FStencils, idpy/LBM/LBM.py:94-185).
struct EStencil {
  int QE;
  const int* E;      // QE*D ints
  const double* EW;  // QE weights
};

// Shan-Chen force at site g (IDShanChenForce.py:85-114, LBM.py:1562-1580):
// F_d = -G psi(x) sum_e w_e e_d psi(x+e)
template <int D>
inline void sc_force(const Geo& geo, const EStencil& es, const double* psi,
                     int64_t g, double G, double* F) {
  int p[3];
  geo.pos(g, p);
  for (int d = 0; d < D; d++) F[d] = 0.;
  for (int k = 0; k < es.QE; k++) {
    int e[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) e[d] = es.E[k * D + d];
    const double np = psi[geo.nb(p, e)];
    for (int d = 0; d < D; d++) F[d] += e[d] * es.EW[k] * np;
  }
  const double lpsi = psi[g];
  for (int d = 0; d < D; d++) F[d] *= -G * lpsi;
}

// Equilibrium + Guo term for population q (LBM.py:1582-1608;
// Equilibria.py:107-138, Forcings.py:104-127 give the same polynomial).
template <class S>
inline void eq_guo(int q, double ln, const double* lu, const double* F,
                   double u_dot_u, double F_dot_u, double* leq, double* lguo) {
  double u_dot_xi = 0., F_dot_xi = 0.;
  for (int d = 0; d < S::D; d++) {
    u_dot_xi += lu[d] * S::xi[q][d];
    F_dot_xi += F[d] * S::xi[q][d];
  }
  double e = 1.;
  e += +u_dot_xi * kCM2 + 0.5 * u_dot_xi * u_dot_xi * kCM4;
  e += -0.5 * u_dot_u * kCM2;
  *leq = e * ln * S::w[q];
  double gu = 0.;
  gu += +F_dot_xi * kCM2 + F_dot_xi * u_dot_xi * kCM4;
  gu += -F_dot_u * kCM2;
  *lguo = gu * S::w[q];
}

// ------------------------------------------------ legacy class (SoA always)
template <class S>
void legacy_step(const Geo& geo, const EStencil& es, double G, double omega,
                 int psi_kind, double*& pop, double*& pop_swap, double* n,
                 double* u, double* psi) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  // K_ComputeMoments (LBM.py:1667-1698)
  for (int64_t g = 0; g < V; g++) {
    double lu[3] = {0., 0., 0.};
    double ln = 0.;
    for (int q = 0; q < Q; q++) {
      const double lpop = pop[g + q * V];
      ln += lpop;
      for (int d = 0; d < D; d++) lu[d] += lpop * S::xi[q][d];
    }
    n[g] = ln;
    for (int d = 0; d < D; d++) u[g + d * V] = lu[d] / ln;
  }
  // K_ComputePsi (LBM.py:1613-1631)
  for (int64_t g = 0; g < V; g++) psi[g] = psi_of(psi_kind, n[g]);
  // K_Collision_ShanChenGuoMultiPhase (LBM.py:1529-1611)
  for (int64_t g = 0; g < V; g++) {
    double F[3];
    sc_force<D>(geo, es, psi, g, G, F);
    const double ln = n[g];
    double lu[3];
    for (int d = 0; d < D; d++) {
      lu[d] = u[g + V * d] + 0.5 * F[d] / ln;
      u[g + V * d] = lu[d];
    }
    double u_dot_u = 0., F_dot_u = 0.;
    for (int d = 0; d < D; d++) {
      u_dot_u += lu[d] * lu[d];
      F_dot_u += F[d] * lu[d];
    }
    for (int q = 0; q < Q; q++) {
      double leq, lguo;
      eq_guo<S>(q, ln, lu, F, u_dot_u, F_dot_u, &leq, &lguo);
      pop[g + q * V] = pop[g + q * V] * (1. - omega) + leq * omega +
                       (1. - 0.5 * omega) * lguo;
    }
  }
  // K_StreamPeriodic, pull (LBM.py:1633-1665)
  for (int64_t g = 0; g < V; g++) {
    int p[3];
    geo.pos(g, p);
    pop_swap[g] = pop[g];
    for (int q = 1; q < Q; q++) {
      int e[3] = {-S::xi[q][0], -S::xi[q][1], -S::xi[q][2]};
      pop_swap[g + q * V] = pop[geo.nb(p, e) + q * V];
    }
  }
  // M_SwapPop (LBM.py:1921-1933)
  double* t = pop;
  pop = pop_swap;
  pop_swap = t;
}

// ------------------------------- legacy-kernel wall loops (SoA, pull stream)
// MainLoopSimpleWalls / MainLoopSimpleWallsGravity (idpy/LBM/MultiPhase.py:
// 351-465, 888-1010): K_ComputeMomentsWalls (LBMKernels.py:220-253, walls ==
// 1), K_ComputePsiWalls (MultiPhaseKernels.py:583-603, walls != 0),
// K_Collision_ShanChenGuoMultiPhaseWalls[Gravity] (MultiPhaseKernels.py:
// 389-560, F[1] += g), K_StreamPeriodic (LBMKernels.py:117-149, every site),
// K_HalfWayBounceBack on the streamed populations (LBMKernels.py:151-184,
// walls == 0, in the serial site order of the CTypes loop), M_SwapPop.
template <class S>
int opposite(int q);

// scf: the Shan-Chen forcing of MainLoopSimpleSC / MainLoopSimpleWallsSC
// (K_Collision_ShanChenMultiPhase[Walls], MultiPhaseKernels.py:132-304) in
// place of the Guo collision. walls == nullptr: no walls (every site fluid,
// no bounce-back pass), as in MainLoopSimpleSC (MultiPhase.py:257-349).
template <class S>
void legacy_walls_step(const Geo& geo, const EStencil& es, double G, double omega,
                       int psi_kind, const uint8_t* walls, double gy, bool scf,
                       double*& pop, double*& pop_swap, double* n, double* u, double* psi) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  auto fluid = [&](int64_t g) { return !walls || walls[g] == 1; };
  for (int64_t g = 0; g < V; g++) {
    if (!fluid(g)) continue;
    double lu[3] = {0., 0., 0.};
    double ln = 0.;
    for (int q = 0; q < Q; q++) {
      const double lpop = pop[g + q * V];
      ln += lpop;
      for (int d = 0; d < D; d++) lu[d] += lpop * S::xi[q][d];
    }
    n[g] = ln;
    for (int d = 0; d < D; d++) u[g + d * V] = lu[d] / ln;
  }
  for (int64_t g = 0; g < V; g++)
    if (!walls || walls[g]) psi[g] = psi_of(psi_kind, n[g]);
  for (int64_t g = 0; g < V; g++) {
    if (scf ? !fluid(g) : (walls && !walls[g])) continue;
    double F[3];
    sc_force<D>(geo, es, psi, g, G, F);
    if (scf) {
      // K_Collision_ShanChenMultiPhase: u_eq = u + F / n / omega, BGK
      // without a forcing term, u <- (u + sum xi f* / n) / 2
      const double ln = n[g];
      double lu[3], lu_post[3] = {0., 0., 0.};
      for (int d = 0; d < D; d++) lu[d] = u[g + V * d] + F[d] / ln / omega;
      double u_dot_u = 0.;
      for (int d = 0; d < D; d++) u_dot_u += lu[d] * lu[d];
      for (int q = 0; q < Q; q++) {
        double u_dot_xi = 0.;
        for (int d = 0; d < D; d++) u_dot_xi += lu[d] * S::xi[q][d];
        double leq = 1.;
        leq += +u_dot_xi * kCM2 + 0.5 * u_dot_xi * u_dot_xi * kCM4;
        leq += -0.5 * u_dot_u * kCM2;
        leq = leq * ln * S::w[q];
        pop[g + q * V] = pop[g + q * V] * (1. - omega) + leq * omega;
        for (int d = 0; d < D; d++) lu_post[d] += pop[g + q * V] * S::xi[q][d];
      }
      for (int d = 0; d < D; d++) u[g + V * d] = 0.5 * (u[g + V * d] + lu_post[d] / ln);
      continue;
    }
    F[1] += gy;
    const double ln = n[g];
    double lu[3];
    for (int d = 0; d < D; d++) {
      lu[d] = u[g + V * d] + 0.5 * F[d] / ln;
      u[g + V * d] = lu[d];
    }
    double u_dot_u = 0., F_dot_u = 0.;
    for (int d = 0; d < D; d++) {
      u_dot_u += lu[d] * lu[d];
      F_dot_u += F[d] * lu[d];
    }
    for (int q = 0; q < Q; q++) {
      double leq, lguo;
      eq_guo<S>(q, ln, lu, F, u_dot_u, F_dot_u, &leq, &lguo);
      pop[g + q * V] = pop[g + q * V] * (1. - omega) + leq * omega +
                       (1. - 0.5 * omega) * lguo;
    }
  }
  for (int64_t g = 0; g < V; g++) {
    int p[3];
    geo.pos(g, p);
    pop_swap[g] = pop[g];
    for (int q = 1; q < Q; q++) {
      int e[3] = {-S::xi[q][0], -S::xi[q][1], -S::xi[q][2]};
      pop_swap[g + q * V] = pop[geo.nb(p, e) + q * V];
    }
  }
  for (int64_t g = 0; walls && g < V; g++) {
    if (walls[g] != 0) continue;
    int p[3];
    geo.pos(g, p);
    pop_swap[g] = 0.;
    for (int q = 1; q < Q; q++) {
      const int qd = opposite<S>(q);
      int e[3] = {S::xi[qd][0], S::xi[qd][1], S::xi[qd][2]};
      pop_swap[geo.nb(p, e) + qd * V] = pop_swap[g + q * V];
      pop_swap[g + q * V] = 0.;
    }
  }
  double* t = pop;
  pop = pop_swap;
  pop_swap = t;
}

// ------------------------------------------- Meta class, BGK or MRT (push)
struct MRTMats {
  const double* A;   // M^-1 (1-S) M     (Q x Q, row-major)
  const double* B;   // M^-1 S M
  const double* Gg;  // M'^-1 (1-S/2) M'
};

// The same MRT collision evaluated in moment space (diagonal S, rates s):
//   f* = f - Mw^-1 [s o (Mw f - Mw f_eq)] + Mp^-1 [(1 - s/2) o (Mp F_raw)]
// written through the lattice monomials (Mw = Tm Mm, Mm[k][q] = xi_q^e_k),
// Mw = Lt Mp and Mw^-1 = W Mw^T diag(Dinv):
//   f* = f + W Mm^T Tm^T b,  b = Dinv o (Lt mg - md),
//   md = s o (Mw f - Mw f_eq),  mg = (1 - s/2) o (Mp F_raw),
// with Mw f_eq = (n, n u_a, n u_a u_b, 0, ...) on the Hermite rows and
// Mp F_raw from the coordinates of F_raw on w o {1, xi_a, xi_a xi_b}. The
// tables come from oracle/hermite_mrt.py (its own sympy bases). Every sum
// starts from its first non-zero term and runs in index order: monomial
// sums over opposite pairs (f_q + f_-q for even monomials, f_q - f_-q for
// odd ones), so that the HIP kernel's moment-space collision, which
// evaluates the same sums in the same order, can be followed closely over
// long runs; the dense path above is the independent check of the algebra.
struct MomTabs {
  int NP;
  const int* pairs;   // first velocity of each opposite pair
  const int* mm;      // Q x Q monomial values in {-1, 0, 1}
  const int* odd;     // Q: parity of each monomial
  const double* tm;   // Q x Q
  const double* lt;   // Q x Q, unit lower-triangular
  const double* dinv; // Q
  const int* ht;      // 3 x Q: Hermite order (0..2, 3 = higher), a, b
  const double* mpv;  // NB x Q: Mp moments of the basis vectors
  const double* s;    // Q relaxation rates
};

struct Sum {
  double v = 0.;
  bool any = false;
  void add(double c, double t) {
    if (c == 0.) return;
    const double x = c == 1. ? t : c == -1. ? -t : c * t;
    v = any ? (c == -1. ? v - t : v + x) : x;
    any = true;
  }
  double get() const { return any ? v : 0.; }
};

// Index of the opposite velocity (IdpyStencil.opposite).
template <class S>
int opposite(int q) {
  for (int k = 0; k < S::Q; k++) {
    bool same = true;
    for (int d = 0; d < 3; d++) same = same && S::xi[k][d] == -S::xi[q][d];
    if (same) return k;
  }
  return q;
}

// K_IsotropyFilter (idpy/LBM/LBMKernelsMeta.py:895-1080) on the nine
// post-collision populations of one D2Q9 site, in place: Hermite moments
// m_k = sum_q M[k][q] f_q normalised by m_0 (:934-956), the xy, xx - yy,
// xxy and xyy moments replaced by their values in terms of the others
// (:962-1015), the populations solved back (f = Ai b, the inverse of the
// system's matrix; the reference expands the symbolic solution instead) and
// rescaled by m_0 (:1062-1072). M and Ai come from oracle/iso_filter.py.
struct IsoMats {
  const double* M;   // 9 x 9 moment rows (1, x, y, xx, xy, yy, xxy, xyy, xxyy)
  const double* Ai;  // 9 x 9 inverse of the filter's system
};
inline void iso_filter_site(const IsoMats& t, double* f) {
  double m[9];
  for (int k = 0; k < 9; k++) {
    double a = 0.;
    for (int q = 0; q < 9; q++) a += t.M[k * 9 + q] * f[q];
    m[k] = a;
  }
  for (int k = 1; k < 9; k++) m[k] /= m[0];
  const double b[9] = {1.,
                       m[1],
                       m[2],
                       m[3],
                       m[1] * m[2],
                       m[1] * m[1] - m[2] * m[2],
                       2. * m[4] * m[1] + (m[3] + 1. / 3) * m[2] - 2. * m[2] * m[1] * m[1],
                       2. * m[4] * m[2] + (m[5] + 1. / 3) * m[1] - 2. * m[1] * m[2] * m[2],
                       m[8]};
  for (int q = 0; q < 9; q++) {
    double a = 0.;
    for (int k = 0; k < 9; k++) a += t.Ai[q * 9 + k] * b[k];
    f[q] = a * m[0];
  }
}

// With iso (D2Q9, BGK) the step of MainLoopIsoFilter (MultiPhase.py:
// 1099-1223): K_ForceCollideSCMPMeta collides in place, K_IsotropyFilter
// filters, K_StreamPeriodicMeta push-streams -- the same as filtering f*
// between the collision and the push of the fused step.
// Sparse form of the moment-space tables: for every sum, its non-zero
// terms in increasing index order (the order the sums are evaluated in).
struct Term {
  int i;
  double c;
};
struct MomSparse {
  std::vector<Term> R[19], Tm[19], Mp[19], Lt[19], TmT[19], E[9], O[9];
  bool col[19];
  int opp[19];
};

template <class S>
MomSparse mom_sparse(const MomTabs& t) {
  constexpr int Q = S::Q, NB = S::D == 2 ? 6 : 10;
  MomSparse m;
  for (int q = 0; q < Q; q++) m.opp[q] = opposite<S>(q);
  for (int k = 0; k < Q; k++) {
    m.col[k] = false;
    for (int b = 0; b < NB; b++)
      if (t.mpv[b * Q + k] != 0.) {
        m.col[k] = true;
        m.Mp[k].push_back({b, t.mpv[b * Q + k]});
      }
    for (int p = 0; p < t.NP; p++)
      if (t.mm[k * Q + t.pairs[p]]) m.R[k].push_back({p, double(t.mm[k * Q + t.pairs[p]])});
    for (int j2 = 0; j2 < Q; j2++)
      if (t.tm[k * Q + j2] != 0.) m.Tm[k].push_back({j2, t.tm[k * Q + j2]});
  }
  for (int k = 0; k < Q; k++)
    for (int j2 = 1; j2 < k; j2++)
      if (m.col[j2] && t.lt[k * Q + j2] != 0.) m.Lt[k].push_back({j2, t.lt[k * Q + j2]});
  for (int j2 = 0; j2 < Q; j2++)
    for (int k = 1; k < Q; k++)
      if (t.tm[k * Q + j2] != 0.) m.TmT[j2].push_back({k, t.tm[k * Q + j2]});
  for (int p = 0; p < t.NP; p++)
    for (int k = 0; k < Q; k++) {
      const int c = t.mm[k * Q + t.pairs[p]];
      if (c) (t.odd[k] ? m.O[p] : m.E[p]).push_back({k, double(c)});
    }
  return m;
}

inline double sum_terms(const std::vector<Term>& ts, const double* x) {
  Sum a;
  for (const Term& e : ts) a.add(e.c, x[e.i]);
  return a.get();
}

template <class S>
void mrt_moment(const MomTabs& t, const MomSparse& sp, const double* f, double n,
                const double* u, const double* F, double* fo) {
  constexpr int D = S::D, Q = S::Q;
  double ps[9], pd[9];
  for (int p = 0; p < t.NP; p++) {
    const int q = t.pairs[p], o = sp.opp[q];
    ps[p] = f[q] + f[o];
    pd[p] = f[q] - f[o];
  }
  double R[Q];
  double dn = 0.;
  for (int q = 0; q < Q; q++) dn += f[q];
  R[0] = dn;
  for (int k = 1; k < Q; k++) R[k] = sum_terms(sp.R[k], t.odd[k] ? pd : ps);
  const double nu[3] = {n * u[0], n * u[1], D == 3 ? n * u[2] : 0.};
  double md[Q];
  for (int k = 1; k < Q; k++) {
    double m = sum_terms(sp.Tm[k], R);
    const int ord = t.ht[k], ha = t.ht[Q + k], hb = t.ht[2 * Q + k];
    if (ord == 1) m = m - nu[ha];
    if (ord == 2) m = m - nu[ha] * u[hb];
    md[k] = t.s[k] * m;
  }
  // coordinates of the raw Guo term on w o {1, xi_a, xi_a xi_b}
  const double ux = u[0], uy = u[1], uz = D == 3 ? u[2] : 0.;
  const double Fx = F[0], Fy = F[1], Fz = D == 3 ? F[2] : 0.;
  double d[10];
  d[0] = -3. * (Fx * ux + Fy * uy + Fz * uz);
  if (D == 2) {
    d[1] = 3. * Fx; d[2] = 3. * Fy;
    d[3] = 9. * Fx * ux; d[4] = 9. * Fy * uy; d[5] = 9. * (Fx * uy + Fy * ux);
  } else {
    d[1] = 3. * Fx; d[2] = 3. * Fy; d[3] = 3. * Fz;
    d[4] = 9. * Fx * ux; d[5] = 9. * Fy * uy; d[6] = 9. * Fz * uz;
    d[7] = 9. * (Fx * uy + Fy * ux);
    d[8] = 9. * (Fx * uz + Fz * ux);
    d[9] = 9. * (Fy * uz + Fz * uy);
  }
  double mg[Q];
  for (int k = 1; k < Q; k++)
    if (sp.col[k]) mg[k] = (1. - 0.5 * t.s[k]) * sum_terms(sp.Mp[k], d);
  double bk[Q];
  for (int k = 1; k < Q; k++) {
    Sum l;
    if (sp.col[k]) l.add(1., mg[k]);
    for (const Term& e : sp.Lt[k]) l.add(e.c, mg[e.i]);
    bk[k] = t.dinv[k] * (l.any ? l.get() - md[k] : -md[k]);
  }
  double cm[Q];
  for (int j = 0; j < Q; j++) cm[j] = sum_terms(sp.TmT[j], bk);
  fo[0] = f[0] + S::w[0] * cm[0];
  for (int p = 0; p < t.NP; p++) {
    const int q = t.pairs[p], o = sp.opp[q];
    const double wE = S::w[q] * sum_terms(sp.E[p], cm), wO = S::w[q] * sum_terms(sp.O[p], cm);
    fo[q] = f[q] + wE + wO;
    fo[o] = f[o] + wE - wO;
  }
}

// Printed sums of the Meta class's generated code (oracle/ref15_terms.py):
// term t of population q, off[q] <= t < off[q+1], is coef[t] times up to
// three variables fac[3t..3t+2] (-1 for none) of
//   v = {ln_0, lu_0_0, lu_0_1, lu_0_2, d_psi_0, d_psi_1, d_psi_2, lpop_q},
// multiplied left to right; the terms are summed left to right, as a C
// compiler evaluates `c0*a*b + c1*c - ...` (idpy/LBM/Collisions.py:251-273,
// idpy/LBM/LBMKernelsMeta.py:282-390).
struct Terms {
  const int* off;
  const double* coef;
  const int* fac;
};
inline double eval_terms(const Terms& t, int q, const double* v) {
  double acc = 0.;
  for (int k = t.off[q]; k < t.off[q + 1]; k++) {
    double p = t.coef[k];
    for (int j = 0; j < 3; j++)
      if (t.fac[3 * k + j] >= 0) p = p * v[t.fac[3 * k + j]];
    acc = k == t.off[q] ? p : acc + p;
  }
  return acc;
}

// The dense collision in the basis-coordinate form the library's tables
// use: f*_q = sum_k A[q][k] f_k + sum_k beta[k][q] c_k + gamma[k][q] d_k with
// c, d the coordinates of f_eq and of the raw Guo term on
// w o {1, xi_a, xi_a xi_b (a <= b)} (2D order {1, x, y, xx, yy, xy}, 3D
// {1, x, y, z, xx, yy, zz, xy, xz, yz}).
struct DenseTabs {
  const double* A;      // Q x Q
  const double* beta;   // NB x Q
  const double* gamma;  // NB x Q
};
template <class S>
void dense_collide(const DenseTabs& t, const double* f, double n, const double* u,
                   const double* F, double* fo) {
  constexpr int D = S::D, Q = S::Q, NB = D == 2 ? 6 : 10;
  double uu = 0., Fu = 0.;
  for (int a = 0; a < D; a++) {
    uu += u[a] * u[a];
    Fu += F[a] * u[a];
  }
  double c[NB], d[NB];
  c[0] = n - 1.5 * n * uu;
  d[0] = -3. * Fu;
  int k = 1;
  for (int a = 0; a < D; a++, k++) {
    c[k] = 3. * n * u[a];
    d[k] = 3. * F[a];
  }
  for (int a = 0; a < D; a++, k++) {
    c[k] = 4.5 * n * u[a] * u[a];
    d[k] = 9. * F[a] * u[a];
  }
  for (int a = 0; a < D; a++)
    for (int b = a + 1; b < D; b++, k++) {
      c[k] = 9. * n * u[a] * u[b];
      d[k] = 9. * (F[a] * u[b] + F[b] * u[a]);
    }
  for (int q = 0; q < Q; q++) {
    double acc = 0.;
    for (int j = 0; j < Q; j++) acc += t.A[q * Q + j] * f[j];
    for (int j = 0; j < NB; j++) acc += t.beta[j * Q + q] * c[j] + t.gamma[j * Q + q] * d[j];
    fo[q] = acc;
  }
}

template <class S>
void meta_step(const Geo& geo, const EStencil& es, const Ord& ord, double G,
               double omega, int psi_kind, const MRTMats* mrt, double*& pop,
               double*& pop_swap, double* n, double* psi, const uint8_t* walls,
               const IsoMats* iso = nullptr, const MomTabs* mom = nullptr,
               const MomSparse* msp = nullptr, const Terms* terms = nullptr,
               const DenseTabs* dense = nullptr) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  // K_ComputeDensityPsiMeta (MultiPhaseKernelsMeta.py:107-167); with walls
  // K_ComputeDensityPsiWallsMeta (:42-104), fluid sites only
  for (int64_t g = 0; g < V; g++) {
    if (walls && walls[g] != 1) continue;
    double ln = 0.;
    for (int q = 0; q < Q; q++) ln += pop[ord(g, q)];
    psi[g] = psi_of(psi_kind, ln);
    n[g] = ln;
  }
  // K_ForceCollideStreamSCMPMeta / _MRTMeta (MultiPhaseKernelsMeta.py:169-345,
  // 891-1075); with walls K_ForceCollideStreamWallsSCMPMeta (:524-716)
  for (int64_t g = 0; g < V; g++) {
    if (walls && walls[g] != 1) continue;
    double F[3] = {0., 0., 0.};
    sc_force<D>(geo, es, psi, g, G, F);
    const double ln = n[g];
    double lpop[Q];
    double lu[3] = {0., 0., 0.};
    for (int q = 0; q < Q; q++) {
      lpop[q] = pop[ord(g, q)];
      for (int d = 0; d < D; d++) lu[d] += lpop[q] * S::xi[q][d];
    }
    // VelocityCode (IdpyStencils.py:444-561) then Guo shift u += F/(2n)
    // (Forcings.py:309-329)
    for (int d = 0; d < D; d++) lu[d] = lu[d] / ln + 0.5 * F[d] / ln;
    double u_dot_u = 0., F_dot_u = 0.;
    for (int d = 0; d < D; d++) {
      u_dot_u += lu[d] * lu[d];
      F_dot_u += F[d] * lu[d];
    }
    double feq[Q], fg[Q], fout[Q];
    for (int q = 0; q < Q; q++) eq_guo<S>(q, ln, lu, F, u_dot_u, F_dot_u, &feq[q], &fg[q]);
    if (dense) {
      dense_collide<S>(*dense, lpop, ln, lu, F, fout);
    } else if (terms) {
      // the printed BGK + Guo sums of K_ForceCollideStreamSCMPMeta
      double v[8] = {ln, lu[0], lu[1], lu[2], F[0], F[1], F[2], 0.};
      for (int q = 0; q < Q; q++) {
        v[7] = lpop[q];
        fout[q] = eval_terms(*terms, q, v);
      }
    } else if (mom) {
      mrt_moment<S>(*mom, *msp, lpop, ln, lu, F, fout);
    } else if (!mrt) {
      // SRT: (1-w) f + w f_eq + (1-w/2) F_guo (Collisions.py:163-185)
      for (int q = 0; q < Q; q++)
        fout[q] = lpop[q] * (1. - omega) + feq[q] * omega + (1. - 0.5 * omega) * fg[q];
    } else {
      for (int q = 0; q < Q; q++) {
        double a = 0.;
        for (int k = 0; k < Q; k++)
          a += mrt->A[q * Q + k] * lpop[k] + mrt->B[q * Q + k] * feq[k] +
               mrt->Gg[q * Q + k] * fg[k];
        fout[q] = a;
      }
    }
    if constexpr (S::Q == 9)
      if (iso) iso_filter_site(*iso, fout);
    // push stream: pop_swap[x + xi_q, q] = f*_q (Collisions.py:487-530);
    // with walls, a destination that is a solid site sends f*_q back to x as
    // the opposite population (Collisions.py:835-850)
    int p[3];
    geo.pos(g, p);
    for (int q = 0; q < Q; q++) {
      int e[3] = {S::xi[q][0], S::xi[q][1], S::xi[q][2]};
      int64_t dst = geo.nb(p, e);
      int qd = q;
      if (walls && q != 0 && walls[dst] == 0) {
        dst = g;
        qd = opposite<S>(q);
      }
      pop_swap[ord(dst, qd)] = fout[q];
    }
  }
  double* t = pop;
  pop = pop_swap;
  pop_swap = t;
}

// ComputeMomentsMeta: n, psi, then u = sum xi f / n + F/(2n)
// (MultiPhase.py:1565-1662, MultiPhaseKernelsMeta.py:1077-1219)
template <class S>
void meta_moments(const Geo& geo, const EStencil& es, const Ord& ordp,
                  bool u_aos, double G, int psi_kind, const double* pop,
                  double* n, double* u, double* psi, const uint8_t* walls) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  // with walls: K_ComputeDensityPsiWallsMeta and
  // K_ComputeVelocityAfterForceSCMPWallsMeta (MultiPhaseKernelsMeta.py:42-104,
  // 1221-1365), fluid sites only
  for (int64_t g = 0; g < V; g++) {
    if (walls && walls[g] != 1) continue;
    double ln = 0.;
    for (int q = 0; q < Q; q++) ln += pop[ordp(g, q)];
    psi[g] = psi_of(psi_kind, ln);
    n[g] = ln;
  }
  for (int64_t g = 0; g < V; g++) {
    if (walls && walls[g] != 1) continue;
    double F[3];
    sc_force<D>(geo, es, psi, g, G, F);
    double lu[3] = {0., 0., 0.};
    for (int q = 0; q < Q; q++) {
      const double f = pop[ordp(g, q)];
      for (int d = 0; d < D; d++) lu[d] += f * S::xi[q][d];
    }
    for (int d = 0; d < D; d++) {
      const double v = lu[d] / n[g] + 0.5 * F[d] / n[g];
      u[u_aos ? d + g * D : g + d * V] = v;
    }
  }
}

// K_InitPopulations: f = f_eq(n, u) order 2 (LBM.py:1700-1743)
template <class S>
void init_pop(const Geo& geo, const Ord& ord, const double* n, const double* u,
              double* pop) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  for (int64_t g = 0; g < V; g++) {
    const double ln = n[g];
    double lu[3] = {0., 0., 0.}, u_dot_u = 0.;
    for (int d = 0; d < D; d++) {
      lu[d] = u[g + d * V];
      u_dot_u += lu[d] * lu[d];
    }
    for (int q = 0; q < Q; q++) {
      double u_dot_xi = 0.;
      for (int d = 0; d < D; d++) u_dot_xi += lu[d] * S::xi[q][d];
      double e = 1.;
      e += u_dot_xi * kCM2;
      e += 0.5 * u_dot_xi * u_dot_xi * kCM4;
      e -= 0.5 * u_dot_u * kCM2;
      pop[ord(g, q)] = e * ln * S::w[q];
    }
  }
}

// K_InitPopulationsMeta with its printed sums (LBMKernelsMeta.py:282-390)
template <class S>
void init_pop_terms(const Geo& geo, const Ord& ord, const double* n, const double* u,
                    const Terms& t, double* pop) {
  constexpr int D = S::D, Q = S::Q;
  const int64_t V = geo.V;
  for (int64_t g = 0; g < V; g++) {
    double v[8] = {n[g], 0., 0., 0., 0., 0., 0., 0.};
    for (int d = 0; d < D; d++) v[1 + d] = u[g + d * V];
    for (int q = 0; q < Q; q++) pop[ord(g, q)] = eval_terms(t, q, v);
  }
}

bool make_geo(int D, const int* L, Geo* geo) {
  if (D < 2 || D > 3) return false;
  geo->D = D;
  geo->L[0] = L[0];
  geo->L[1] = L[1];
  geo->L[2] = D > 2 ? L[2] : 1;
  geo->V = (int64_t)geo->L[0] * geo->L[1] * geo->L[2];
  for (int d = 0; d < D; d++)
    if (L[d] <= 0) return false;
  return true;
}

}  // namespace

extern "C" {

// The function that psi kind 3 calls (double psi(double n)), or NULL.
int cpuref_set_psi_fn(void* fn) {
  g_psi_fn = reinterpret_cast<double (*)(double)>(fn);
  return 0;
}

// Velocity set ids: 0 = D2Q9, 1 = D3Q19.
// Runs nsteps of the legacy class MainLoop (SoA). After the call pop holds
// the state (the function copies it back when the swap count is odd); n, u,
// psi hold the values computed during the last step, as on the reference.
int cpuref_legacy_run(int xi_id, const int* L, int QE, const int* E,
                      const double* EW, double G, double omega, int psi_kind,
                      double* pop, double* pop_swap, double* n, double* u,
                      double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  EStencil es{QE, E, EW};
  double *a = pop, *b = pop_swap;
  for (int s = 0; s < nsteps; s++) {
    if (xi_id == 0) legacy_step<SD2Q9>(geo, es, G, omega, psi_kind, a, b, n, u, psi);
    else legacy_step<SD3Q19>(geo, es, G, omega, psi_kind, a, b, n, u, psi);
  }
  const int Q = xi_id == 0 ? 9 : 19;
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// Runs nsteps of MainLoopSimpleWalls (gy == 0) or MainLoopSimpleWallsGravity
// (SoA), or with scf != 0 MainLoopSimpleWallsSC (MainLoopSimpleSC when walls
// is NULL); walls is V flags, 1 fluid / 0 solid. pop holds the state after
// the call; n, u, psi hold the values of the last step.
int cpuref_legacy_walls_run(int xi_id, const int* L, int QE, const int* E,
                            const double* EW, double G, double omega, int psi_kind,
                            const uint8_t* walls, double gy, int scf, double* pop,
                            double* pop_swap, double* n, double* u, double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  EStencil es{QE, E, EW};
  double *a = pop, *b = pop_swap;
  for (int s = 0; s < nsteps; s++) {
    if (xi_id == 0)
      legacy_walls_step<SD2Q9>(geo, es, G, omega, psi_kind, walls, gy, scf != 0, a, b, n, u,
                               psi);
    else
      legacy_walls_step<SD3Q19>(geo, es, G, omega, psi_kind, walls, gy, scf != 0, a, b, n, u,
                                psi);
  }
  const int Q = xi_id == 0 ? 9 : 19;
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// Runs nsteps of the Meta class MainLoop (mrt == 0) or MainLoopMRT (A, B, Gg
// given); with walls (V flags, 1 fluid / 0 solid; NULL for none) the step of
// MainLoopWalls. aos != 0 selects the CTypes 'cpu' ordering pop[q + g*Q].
int cpuref_meta_run(int xi_id, const int* L, int QE, const int* E,
                    const double* EW, double G, double omega, int psi_kind,
                    int aos, const double* A, const double* B, const double* Gg,
                    double* pop, double* pop_swap, double* n, double* psi,
                    int nsteps, const uint8_t* walls) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, Q};
  MRTMats m{A, B, Gg};
  const MRTMats* mp = A ? &m : nullptr;
  double *a = pop, *b = pop_swap;
  for (int s = 0; s < nsteps; s++) {
    if (xi_id == 0) meta_step<SD2Q9>(geo, es, ord, G, omega, psi_kind, mp, a, b, n, psi, walls);
    else meta_step<SD3Q19>(geo, es, ord, G, omega, psi_kind, mp, a, b, n, psi, walls);
  }
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// nsteps of the Meta class MainLoop with the collision given as its printed
// sums (Terms above; oracle/ref15_terms.py builds them).
int cpuref_meta_run_terms(int xi_id, const int* L, int QE, const int* E, const double* EW,
                          double G, int psi_kind, int aos, const int* off, const double* coef,
                          const int* fac, double* pop, double* pop_swap, double* n,
                          double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, Q};
  Terms t{off, coef, fac};
  double *a = pop, *b = pop_swap;
  for (int k = 0; k < nsteps; k++) {
    if (xi_id == 0)
      meta_step<SD2Q9>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr,
                       nullptr, nullptr, &t);
    else
      meta_step<SD3Q19>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr,
                        nullptr, nullptr, &t);
  }
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// nsteps of the Meta class MainLoop with the dense basis-coordinate tables
// (DenseTabs above).
int cpuref_meta_run_dense(int xi_id, const int* L, int QE, const int* E, const double* EW,
                          double G, int psi_kind, int aos, const double* A, const double* beta,
                          const double* gamma, double* pop, double* pop_swap, double* n,
                          double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, Q};
  DenseTabs t{A, beta, gamma};
  double *a = pop, *b = pop_swap;
  for (int k = 0; k < nsteps; k++) {
    if (xi_id == 0)
      meta_step<SD2Q9>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr,
                       nullptr, nullptr, nullptr, &t);
    else
      meta_step<SD3Q19>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr,
                        nullptr, nullptr, nullptr, &t);
  }
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// K_InitPopulationsMeta with printed sums (u in SoA order, u[g + d V]).
int cpuref_init_pop_terms(int xi_id, const int* L, int aos, const double* n, const double* u,
                          const int* off, const double* coef, const int* fac, double* pop) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  Ord ord{aos != 0, geo.V, Q};
  Terms t{off, coef, fac};
  if (xi_id == 0) init_pop_terms<SD2Q9>(geo, ord, n, u, t, pop);
  else init_pop_terms<SD3Q19>(geo, ord, n, u, t, pop);
  return 0;
}

// nsteps of MainLoopMRT with a diagonal relaxation matrix, the collision
// evaluated in moment space (mrt_moment above). Table layout as MomTabs.
int cpuref_meta_run_mom(int xi_id, const int* L, int QE, const int* E, const double* EW,
                        double G, int psi_kind, int aos, int NP, const int* pairs,
                        const int* mm, const int* odd, const double* tm, const double* lt,
                        const double* dinv, const int* ht, const double* mpv, const double* s,
                        double* pop, double* pop_swap, double* n, double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  if (NP != (Q - 1) / 2) return -1;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, Q};
  MomTabs t{NP, pairs, mm, odd, tm, lt, dinv, ht, mpv, s};
  const MomSparse sp = xi_id == 0 ? mom_sparse<SD2Q9>(t) : mom_sparse<SD3Q19>(t);
  double *a = pop, *b = pop_swap;
  for (int k = 0; k < nsteps; k++) {
    if (xi_id == 0)
      meta_step<SD2Q9>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr, &t,
                       &sp);
    else
      meta_step<SD3Q19>(geo, es, ord, G, 1., psi_kind, nullptr, a, b, n, psi, nullptr, nullptr,
                        &t, &sp);
  }
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * Q);
  return 0;
}

// nsteps of MainLoopIsoFilter (D2Q9, BGK, push): the Meta step with the
// isotropy filter between collision and stream. M, Ai: oracle/iso_filter.py.
int cpuref_iso_run(const int* L, int QE, const int* E, const double* EW, double G,
                   double omega, int psi_kind, int aos, const double* M, const double* Ai,
                   double* pop, double* pop_swap, double* n, double* psi, int nsteps) {
  Geo geo;
  if (!make_geo(2, L, &geo)) return -1;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, 9};
  IsoMats t{M, Ai};
  double *a = pop, *b = pop_swap;
  for (int s = 0; s < nsteps; s++)
    meta_step<SD2Q9>(geo, es, ord, G, omega, psi_kind, nullptr, a, b, n, psi, nullptr, &t);
  if (a != pop) memcpy(pop, a, sizeof(double) * geo.V * 9);
  return 0;
}

// The filter alone on nsites independent sites, f[q + 9 s], in place.
int cpuref_iso_filter(int64_t nsites, const double* M, const double* Ai, double* f) {
  IsoMats t{M, Ai};
  for (int64_t s = 0; s < nsites; s++) iso_filter_site(t, f + 9 * s);
  return 0;
}

int cpuref_meta_moments(int xi_id, const int* L, int QE, const int* E,
                        const double* EW, double G, int psi_kind, int aos,
                        const double* pop, double* n, double* u, double* psi,
                        const uint8_t* walls) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  EStencil es{QE, E, EW};
  Ord ord{aos != 0, geo.V, Q};
  if (xi_id == 0)
    meta_moments<SD2Q9>(geo, es, ord, aos != 0, G, psi_kind, pop, n, u, psi, walls);
  else
    meta_moments<SD3Q19>(geo, es, ord, aos != 0, G, psi_kind, pop, n, u, psi, walls);
  return 0;
}

int cpuref_init_pop(int xi_id, const int* L, int aos, const double* n,
                    const double* u, double* pop) {
  Geo geo;
  if (!make_geo(xi_id == 0 ? 2 : 3, L, &geo)) return -1;
  const int Q = xi_id == 0 ? 9 : 19;
  Ord ord{aos != 0, geo.V, Q};
  if (xi_id == 0) init_pop<SD2Q9>(geo, ord, n, u, pop);
  else init_pop<SD3Q19>(geo, ord, n, u, pop);
  return 0;
}

// Legacy K_InitRadialInterface: n = (ng+nl)/2 - (nl-ng)/2 tanh((r-R)/W),
// centre L//2 (LBM.py:1786-1826, 422-423; MultiPhaseKernels.py:99-112).
int cpuref_init_radial(int D, const int* L, double n_g, double n_l, double R,
                       double W, double* n, double* u) {
  Geo geo;
  if (!make_geo(D, L, &geo)) return -1;
  int c[3] = {geo.L[0] / 2, geo.L[1] / 2, geo.L[2] / 2};
  const double delta_n = n_l - n_g;
  for (int64_t g = 0; g < geo.V; g++) {
    int p[3];
    geo.pos(g, p);
    double dist = 0;
    for (int d = 0; d < D; d++) dist += (c[d] - p[d]) * (c[d] - p[d]);
    const double r = sqrt(dist);
    n[g] = 0.5 * (n_g + n_l) - 0.5 * delta_n * tanh((r - R) / W);
    for (int d = 0; d < D; d++) u[g + d * geo.V] = 0.;
  }
  return 0;
}

// K_InitFlatInterface (idpy/LBM/LBM.py:1745-1784) with
// F_NFlatProfilePeriodic (idpy/LBM/MultiPhaseKernels.py:55-74); u = 0.
int cpuref_init_flat(int D, const int* L, double n_g, double n_l, double width, int dir,
                     int full_flag, double* n, double* u) {
  Geo geo;
  if (!make_geo(D, L, &geo) || dir < 0 || dir >= D) return -1;
  const int c = geo.L[dir] / 2, size = geo.L[dir];
  const double delta_n = full_flag * (n_l - n_g) + (1 - full_flag) * (n_g - n_l);
  for (int64_t g = 0; g < geo.V; g++) {
    int p[3];
    geo.pos(g, p);
    const int x = p[dir], xm = size / 2;
    int xp = 0;
    const int delta = c - xm;
    if (delta >= 0) xp = x < c - xm ? x - (c - xm) + size : x - (c - xm);
    else xp = x < c + xm ? x - (c + xm) + size : x - (c + xm);
    const double prof =
        tanh((double)(xp - (xm - 0.5 * width))) - tanh((double)(xp - (xm + 0.5 * width)));
    n[g] = 0.5 * (n_g + n_l) + 0.5 * delta_n * (prof - 1.);
    for (int d = 0; d < D; d++) u[g + d * geo.V] = 0.;
  }
  return 0;
}

// K_InitCylinderInterface (idpy/LBM/LBM.py:1828-1874); u = 0.
int cpuref_init_cylinder(int D, const int* L, double n_g, double n_l, double R, int dir,
                         int full_flag, double* n, double* u) {
  Geo geo;
  if (!make_geo(D, L, &geo) || dir < 0 || dir >= D) return -1;
  int c[3] = {geo.L[0] / 2, geo.L[1] / 2, geo.L[2] / 2};
  const double delta_n = full_flag * (n_l - n_g) + (1 - full_flag) * (n_g - n_l);
  for (int64_t g = 0; g < geo.V; g++) {
    int p[3];
    geo.pos(g, p);
    p[dir] = 0;
    int cp[3] = {c[0], c[1], c[2]};
    cp[dir] = 0;
    double dist = 0;
    for (int d = 0; d < D; d++) dist += (cp[d] - p[d]) * (cp[d] - p[d]);
    n[g] = 0.5 * (n_g + n_l) - 0.5 * delta_n * tanh(sqrt(dist) - R);
    for (int d = 0; d < D; d++) u[g + d * geo.V] = 0.;
  }
  return 0;
}

// K_InitSingleFlatInterface (idpy/LBM/MultiPhaseKernels.py:650-684) with
// F_NSingleFlatProfile (:45-53); u = 0.
int cpuref_init_single_flat(int D, const int* L, double n_g, double n_l, int dir, int full_flag,
                            double* n, double* u) {
  Geo geo;
  if (!make_geo(D, L, &geo) || dir < 0 || dir >= D) return -1;
  const int c = geo.L[dir] / 2;
  const double delta_n = full_flag * (n_l - n_g) + (1 - full_flag) * (n_g - n_l);
  for (int64_t g = 0; g < geo.V; g++) {
    int p[3];
    geo.pos(g, p);
    n[g] = n_g + delta_n * ((tanh((double)(p[dir] - c)) + 1.) * 0.5);
    for (int d = 0; d < D; d++) u[g + d * geo.V] = 0.;
  }
  return 0;
}

}  // extern "C"
