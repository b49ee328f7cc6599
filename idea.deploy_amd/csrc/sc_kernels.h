// sc_kernels.h -- HIP kernels of the Shan-Chen multiphase LBM step for gfx950.
//
// One time step (idpy/LBM/MultiPhase.py:1015-1097):
//   n = sum_q f_q, psi = psi(n)                 K_ComputeDensityPsiMeta
//   F = -G psi(x) sum_e w_e e psi(x+e)          IDShanChenForce.py:85-114
//   u = sum_q xi_q f_q / n + F / (2n)           VelocityCode + Guo shift
//   f*_q = BGK or MRT collision with Guo forcing (Collisions.py:163-185,
//          96-160; Forcings.py:104-181)
//   pop_swap[x + xi_q, q] = f*_q                push streaming
//
// Storage: SoA pop[q*V + g], g = x + L0*(y + L1*z). T = double, or float with
// the shift f~_q = f_q - w_q*rho0 (SHIFT): every linear step of the update
// commutes with the shift, so only the density needs rho0 added back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice_tables.h"

namespace idmx {

// 1/c2 and 1/c2^2 for c2 = 1/3 evaluated in double (CM2, CM4 of
// idpy/LBM/LBM.py:542-569): exactly 3 and 9.
constexpr double kCM2 = 3.0;
constexpr double kCM4 = 9.0;

struct Geom {
  int L0, L1, L2;  // L2 = 1 for two dimensions
  int64_t V;
};

// Number of equilibrium / forcing basis vectors w∘{1, xi_a, xi_a xi_b}.
template <int D> struct NBasis { static constexpr int value = D == 2 ? 6 : 10; };

// A pseudopotential given as a C expression in n (the psi_code kwarg),
// compiled on the host into a stack program (idmx/LBM/psi_code.py): PSI_C
// pushes the next constant, PSI_N the density, the others pop their
// operands and push the result; at most kPsiStack values are live.
enum PsiOp : unsigned char {
  PSI_N = 0, PSI_C, PSI_ADD, PSI_SUB, PSI_MUL, PSI_DIV, PSI_NEG, PSI_EXP, PSI_LOG,
  PSI_SQRT, PSI_POW, PSI_FABS, PSI_TANH, PSI_SINH, PSI_COSH, PSI_SIN, PSI_COS,
  PSI_FMAX, PSI_FMIN, PSI_F32, PSI_NOPS
};
constexpr int kPsiOps = 48, kPsiConsts = 16, kPsiStack = 8;
struct PsiProg {
  int n = 0;                   // number of ops
  unsigned char op[kPsiOps];
  double c[kPsiConsts];
};

template <class T>
struct Phys {
  T G;       // Shan-Chen coupling SC_G
  T omega;   // 1/tau
  T rho0;    // population shift (0 unless SHIFT)
  int psi_kind;
  const T* mrtA;  // Q x Q     M^-1 (1-S) M
  const T* mrtB;  // NB x Q    M^-1 S M v_k
  const T* mrtG;  // NB x Q    M'^-1 (1-S/2) M' v_k
  const PsiProg* prog;  // psi_kind 3: the program (device copy)
};

__device__ inline double dexp(double v) { return exp(v); }
__device__ inline float dexp(float v) { return expf(v); }

// psi(n) for the closed-form pseudopotentials (psi_code kwarg).
template <class T>
__device__ inline T psi_of(int kind, T n) {
  if (kind == 0) return T(1) - dexp(-n);
  if (kind == 1) return dexp(T(-1) / n);
  return n;
}

// Evaluates a psi program at n in FP64 (C's double arithmetic; PSI_F32
// rounds to float as a (float) cast or an f-suffixed literal does). The
// program and the stack depth are uniform over the wave.
__device__ inline double psi_prog(const PsiProg* __restrict__ p, double n) {
  double st[kPsiStack];
  int sp = 0, ci = 0;
  const int len = p->n;
  for (int i = 0; i < len; i++) {
    const int op = p->op[i];
    if (op == PSI_N) { st[sp++] = n; continue; }
    if (op == PSI_C) { st[sp++] = p->c[ci++]; continue; }
    const double b = st[sp - 1];
    switch (op) {
      case PSI_NEG: st[sp - 1] = -b; continue;
      case PSI_EXP: st[sp - 1] = exp(b); continue;
      case PSI_LOG: st[sp - 1] = log(b); continue;
      case PSI_SQRT: st[sp - 1] = sqrt(b); continue;
      case PSI_FABS: st[sp - 1] = fabs(b); continue;
      case PSI_TANH: st[sp - 1] = tanh(b); continue;
      case PSI_SINH: st[sp - 1] = sinh(b); continue;
      case PSI_COSH: st[sp - 1] = cosh(b); continue;
      case PSI_SIN: st[sp - 1] = sin(b); continue;
      case PSI_COS: st[sp - 1] = cos(b); continue;
      case PSI_F32: st[sp - 1] = (double)(float)b; continue;
      default: break;
    }
    const double a = st[sp - 2];
    double r;
    switch (op) {
      case PSI_ADD: r = a + b; break;
      case PSI_SUB: r = a - b; break;
      case PSI_MUL: r = a * b; break;
      case PSI_DIV: r = a / b; break;
      case PSI_POW: r = pow(a, b); break;
      case PSI_FMAX: r = fmax(a, b); break;
      default: r = fmin(a, b); break;
    }
    st[sp - 2] = r;
    sp--;
  }
  return st[0];
}


// Periodic wrap of a coordinate that is at most one period out of range.
__device__ inline int wrap(int v, int L) {
  v = v < 0 ? v + L : v;
  return v >= L ? v - L : v;
}

// Coordinates of the equilibrium (c) and of the raw Guo term (d) on the
// basis v_k = w * {1, xi_a, xi_a xi_b (a <= b)}: 2D {1, x, y, xx, yy, xy},
// 3D {1, x, y, z, xx, yy, zz, xy, xz, yz}. With SHIFT the constant term uses
// sum f~ (= n - rho0), i.e. the coordinates of f_eq - w rho0.
template <class S, class T, bool SHIFT>
__device__ inline void basis_coords(T dn, T n, T ux, T uy, T uz, T Fx, T Fy, T Fz, T uu,
                                    T Fu, T* c, T* d) {
  const T n0 = SHIFT ? dn : n;
  c[0] = n0 - n * T(0.5 * kCM2) * uu;
  d[0] = -T(kCM2) * Fu;
  if (S::D == 2) {
    c[1] = n * T(kCM2) * ux;  c[2] = n * T(kCM2) * uy;
    c[3] = n * T(0.5 * kCM4) * ux * ux;
    c[4] = n * T(0.5 * kCM4) * uy * uy;
    c[5] = n * T(kCM4) * ux * uy;
    d[1] = T(kCM2) * Fx;  d[2] = T(kCM2) * Fy;
    d[3] = T(kCM4) * Fx * ux;
    d[4] = T(kCM4) * Fy * uy;
    d[5] = T(kCM4) * (Fx * uy + Fy * ux);
  } else {
    c[1] = n * T(kCM2) * ux;  c[2] = n * T(kCM2) * uy;  c[3] = n * T(kCM2) * uz;
    c[4] = n * T(0.5 * kCM4) * ux * ux;
    c[5] = n * T(0.5 * kCM4) * uy * uy;
    c[6] = n * T(0.5 * kCM4) * uz * uz;
    c[7] = n * T(kCM4) * ux * uy;
    c[8] = n * T(kCM4) * ux * uz;
    c[9] = n * T(kCM4) * uy * uz;
    d[1] = T(kCM2) * Fx;  d[2] = T(kCM2) * Fy;  d[3] = T(kCM2) * Fz;
    d[4] = T(kCM4) * Fx * ux;
    d[5] = T(kCM4) * Fy * uy;
    d[6] = T(kCM4) * Fz * uz;
    d[7] = T(kCM4) * (Fx * uy + Fy * ux);
    d[8] = T(kCM4) * (Fx * uz + Fz * ux);
    d[9] = T(kCM4) * (Fy * uz + Fz * uy);
  }
}

// Index of the velocity opposite to c_q.
template <class S>
__device__ constexpr int opposite(int q) {
  for (int k = 0; k < S::Q; k++)
    if (S::cx[k] == -S::cx[q] && S::cy[k] == -S::cy[q] && S::cz[k] == -S::cz[q]) return k;
  return q;
}

// Sum of terms with compile-time coefficients, started from its first
// non-zero term (0 + a is not folded to a: signed zeros), coefficients +-1
// applied as additions and subtractions. With the loops unrolled every
// branch on `any` and on the coefficient is resolved at compile time.
template <class T>
struct Acc {
  T v = T(0);
  bool any = false;
  __device__ __forceinline__ void add(T t) {
    v = any ? v + t : t;
    any = true;
  }
  __device__ __forceinline__ void sub(T t) {
    v = any ? v - t : -t;
    any = true;
  }
  __device__ __forceinline__ void addc(double c, T t) {
    if (c == 1.0) add(t);
    else if (c == -1.0) sub(t);
    else if (c != 0.0) add(T(c) * t);
  }
  __device__ __forceinline__ T get() const { return any ? v : T(0); }
};

// True when moment k of some forcing basis vector is non-zero (Mp v_b).
template <class S>
__device__ constexpr bool mpv_row_nz(int k) {
  for (int b = 0; b < NBasis<S::D>::value; b++)
    if (S::MpV[b][k] != 0.0) return true;
  return false;
}

// --------------------------------------------------------------------------
// Isotropy filter of one D2Q9 site's post-collision populations, in place
// (K_IsotropyFilter, idpy/LBM/LBMKernelsMeta.py:895-1080): Hermite moments
// m_k = sum_q IsoM[k][q] f_q (rows 1, x, y, xx, xy, yy, xxy, xyy, xxyy)
// normalised by m_0; the raw xy, xx - yy, xxy and xyy sums set from the
// other moments (:1003-1015); the populations solved back as IsoAi b (IsoAi
// the inverse of the system's matrix) and rescaled by m_0.
template <class S, class T>
__device__ inline void iso_filter(T (&f)[S::Q]) {
  static_assert(S::Q == 9 && S::D == 2, "the isotropy filter is defined for D2Q9");
  T m[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    T a = 0;
#pragma unroll
    for (int q = 0; q < 9; q++)
      if (S::IsoM[k][q] != 0.0) a += T(S::IsoM[k][q]) * f[q];
    m[k] = a;
  }
#pragma unroll
  for (int k = 1; k < 9; k++) m[k] = m[k] / m[0];
  const T third = T(1) / T(3);
  const T b[9] = {T(1),
                  m[1],
                  m[2],
                  m[3],
                  m[1] * m[2],
                  m[1] * m[1] - m[2] * m[2],
                  T(2) * m[4] * m[1] + (m[3] + third) * m[2] - T(2) * m[2] * m[1] * m[1],
                  T(2) * m[4] * m[2] + (m[5] + third) * m[1] - T(2) * m[1] * m[2] * m[2],
                  m[8]};
#pragma unroll
  for (int q = 0; q < 9; q++) {
    T a = 0;
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (S::IsoAi[q][k] != 0.0) a += T(S::IsoAi[q][k]) * b[k];
    f[q] = a * m[0];
  }
}

// --------------------------------------------------------------------------
// Collision of one site. f holds the (shifted when SHIFT) populations, dn is
// sum_q f_q of those values, (gx, gy, gz) the isotropic gradient of psi and
// lpsi = psi(x). Returns f* in fo and the Guo-shifted velocity in u. With BF
// a body force (bx, by, bz) is added to the Shan-Chen force (the gravity of
// K_Collision_ShanChenGuoMultiPhaseWallsGravity, F[1] += g,
// idpy/LBM/MultiPhaseKernels.py:474-560).
template <class S, class T, bool SHIFT, int COLL, bool BF = false>
__device__ inline void collide(const T (&f)[S::Q], T dn, T gx, T gy, T gz,
                               T lpsi, const Phys<T>& P, T (&fo)[S::Q],
                               T (&u)[3], T bx = T(0), T by = T(0), T bz = T(0)) {
  constexpr int Q = S::Q;
  T jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if (S::cx[q] == 1) jx += f[q];
    if (S::cx[q] == -1) jx -= f[q];
    if (S::cy[q] == 1) jy += f[q];
    if (S::cy[q] == -1) jy -= f[q];
    if (S::cz[q] == 1) jz += f[q];
    if (S::cz[q] == -1) jz -= f[q];
  }
  const T n = SHIFT ? P.rho0 + dn : dn;
  const T mGpsi = -P.G * lpsi;
  T Fx = mGpsi * gx, Fy = mGpsi * gy, Fz = S::D == 3 ? mGpsi * gz : T(0);
  if constexpr (BF) {
    Fx += bx;
    Fy += by;
    if (S::D == 3) Fz += bz;
  }
  if constexpr (COLL == 3) {
    // BGK with the Shan-Chen forcing (K_Collision_ShanChenMultiPhase,
    // idpy/LBM/MultiPhaseKernels.py:132-216): the equilibrium velocity is
    // u + F/(n omega), no Guo term, and the velocity written is the mean of
    // the pre- and post-collision velocities.
    const T om = P.omega;
    const T u0x = jx / n, u0y = jy / n, u0z = S::D == 3 ? jz / n : T(0);
    const T lux = u0x + Fx / n / om, luy = u0y + Fy / n / om;
    const T luz = S::D == 3 ? u0z + Fz / n / om : T(0);
    const T uu = lux * lux + luy * luy + luz * luz;
    T px = 0, py = 0, pz = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const T xu = S::cx[q] * lux + S::cy[q] * luy + S::cz[q] * luz;
      T e = T(1);
      e += xu * T(kCM2) + T(0.5) * xu * xu * T(kCM4);
      e += T(-0.5) * uu * T(kCM2);
      T feq = e * n * T(S::w[q]);
      if (SHIFT) feq -= T(S::w[q]) * P.rho0;
      fo[q] = f[q] * (T(1) - om) + feq * om;
      if (S::cx[q] == 1) px += fo[q];
      if (S::cx[q] == -1) px -= fo[q];
      if (S::cy[q] == 1) py += fo[q];
      if (S::cy[q] == -1) py -= fo[q];
      if (S::cz[q] == 1) pz += fo[q];
      if (S::cz[q] == -1) pz -= fo[q];
    }
    u[0] = T(0.5) * (u0x + px / n);
    u[1] = T(0.5) * (u0y + py / n);
    u[2] = S::D == 3 ? T(0.5) * (u0z + pz / n) : T(0);
    return;
  }
  const T inv_n = T(1) / n;
  const T ux = (jx + T(0.5) * Fx) * inv_n;
  const T uy = (jy + T(0.5) * Fy) * inv_n;
  const T uz = S::D == 3 ? (jz + T(0.5) * Fz) * inv_n : T(0);
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
  const T uu = ux * ux + uy * uy + uz * uz;
  const T Fu = Fx * ux + Fy * uy + Fz * uz;
  if (COLL == 0 || COLL == 4) {
    const T om = P.omega;
    const T c_f = T(1) - om;
    const T c_g = T(1) - T(0.5) * om;
    // n-part of f_eq without the shift: SHIFT subtracts w*rho0 from f_eq,
    // i.e. uses dn in place of n for the constant term.
    const T n0 = SHIFT ? dn : n;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const T xu = S::cx[q] * ux + S::cy[q] * uy + S::cz[q] * uz;
      const T xF = S::cx[q] * Fx + S::cy[q] * Fy + S::cz[q] * Fz;
      const T w = T(S::w[q]);
      const T feq = w * (n0 + n * (T(kCM2) * xu + T(0.5 * kCM4) * xu * xu - T(0.5 * kCM2) * uu));
      const T fg = w * (T(kCM2) * xF + T(kCM4) * xF * xu - T(kCM2) * Fu);
      fo[q] = c_f * f[q] + om * feq + c_g * fg;
    }
    // COLL 4: BGK with Guo forcing, then the isotropy filter (D2Q9, FP64)
    if constexpr (COLL == 4) iso_filter<S, T>(fo);
  } else if (COLL == 2) {
    // Diagonal S in moment space:
    //   f* = f - Mw^-1 [s o (Mw f - Mw f_eq)] + Mp^-1 [(1 - s/2) o (Mp F)]
    // (F the raw Guo term; P.mrtA holds the Q rates s), evaluated through
    // the lattice monomials: Mw = Tm Mm with Mm[k][q] = xi_q^e_k in
    // {-1, 0, 1}, Mw = Lt Mp, Mw^-1 = W Mw^T diag(Dinv). Then
    //   f* = f + W Mm^T Tm^T b,  b = Dinv o (Lt mg - md),
    // md = s o (Mw f - Mw f_eq), mg = (1 - s/2) o Mp F. The monomial sums
    // run over opposite pairs (even monomials take f_q + f_-q, odd ones
    // f_q - f_-q), so each moment is a few additions. Row 0 (the mass) is
    // not relaxed: Mw f and Mw f_eq agree there, as do Mp F and 0.
    constexpr int NP = S::NPAIR;
    const T* s = P.mrtA;
    T ps[NP], pd[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {
      ps[p] = f[S::PairQ[p]] + f[S::opp[S::PairQ[p]]];
      pd[p] = f[S::PairQ[p]] - f[S::opp[S::PairQ[p]]];
    }
    // monomial moments R_k = sum_q xi_q^e_k f_q; R_0 = sum_q f_q (shifted
    // populations give the same W-orthogonal moments for k >= 1)
    T R[Q];
    R[0] = dn;
#pragma unroll
    for (int k = 1; k < Q; k++) {
      Acc<T> a;
#pragma unroll
      for (int p = 0; p < NP; p++)
        a.addc(S::Mm[k][S::PairQ[p]], S::MonoOdd[k] ? pd[p] : ps[p]);
      R[k] = a.get();
    }
    // non-equilibrium W-orthogonal moments; Mw f_eq is n, n u_a, n u_a u_b
    // on the rows up to second order and 0 above
    const T nu[3] = {n * ux, n * uy, n * uz};
    const T uv[3] = {ux, uy, uz};
    T md[Q];
#pragma unroll
    for (int k = 1; k < Q; k++) {
      Acc<T> a;
#pragma unroll
      for (int j = 0; j < Q; j++) a.addc(S::Tm[k][j], R[j]);
      T m = a.get();
      if (S::HTord[k] == 1) m = m - nu[S::HTa[k]];
      if (S::HTord[k] == 2) m = m - nu[S::HTa[k]] * uv[S::HTb[k]];
      md[k] = s[k] * m;
    }
    // Mp F from the coordinates d of the raw Guo term on w o {1, xi, xi xi}
    constexpr int NB = NBasis<S::D>::value;
    T cb[NB], d[NB];
    basis_coords<S, T, SHIFT>(dn, n, ux, uy, uz, Fx, Fy, Fz, uu, Fu, cb, d);
    T mg[Q];
#pragma unroll
    for (int k = 1; k < Q; k++) {
      if (!mpv_row_nz<S>(k)) continue;
      Acc<T> a;
#pragma unroll
      for (int b = 0; b < NB; b++) a.addc(S::MpV[b][k], d[b]);
      mg[k] = (T(1) - T(0.5) * s[k]) * a.get();
    }
    T bk[Q];
#pragma unroll
    for (int k = 1; k < Q; k++) {
      Acc<T> l;
      if (mpv_row_nz<S>(k)) l.add(mg[k]);
#pragma unroll
      for (int j = 1; j < k; j++)
        if (mpv_row_nz<S>(j)) l.addc(S::Lt[k][j], mg[j]);
      bk[k] = T(S::Dinv[k]) * (l.any ? l.get() - md[k] : -md[k]);
    }
    // c = Tm^T b, then f*_q = f_q + w_q sum_k Mm[k][q] c_k pair by pair
    T cm[Q];
#pragma unroll
    for (int j = 0; j < Q; j++) {
      Acc<T> a;
#pragma unroll
      for (int k = 1; k < Q; k++) a.addc(S::Tm[k][j], bk[k]);
      cm[j] = a.get();
    }
    fo[0] = f[0] + T(S::w[0]) * cm[0];
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const int q = S::PairQ[p], o = S::opp[S::PairQ[p]];
      Acc<T> E, O;
#pragma unroll
      for (int k = 0; k < Q; k++) {
        if (S::MonoOdd[k]) O.addc(S::Mm[k][q], cm[k]);
        else E.addc(S::Mm[k][q], cm[k]);
      }
      const T wE = T(S::w[q]) * E.get(), wO = T(S::w[q]) * O.get();
      fo[q] = f[q] + wE + wO;
      fo[o] = f[o] + wE - wO;
    }
  } else {
    // f* = A f + sum_k c_k beta_k + sum_k d_k gamma_k, with c_k / d_k the
    // coordinates of f_eq / raw Guo term on w∘{1, xi_a, xi_a xi_b}.
    constexpr int NB = NBasis<S::D>::value;
    T c[NB], d[NB];
    basis_coords<S, T, SHIFT>(dn, n, ux, uy, uz, Fx, Fy, Fz, uu, Fu, c, d);
#pragma unroll
    for (int q = 0; q < Q; q++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < Q; k++) a += P.mrtA[q * Q + k] * f[k];
#pragma unroll
      for (int k = 0; k < NB; k++) a += P.mrtB[k * Q + q] * c[k] + P.mrtG[k * Q + q] * d[k];
      fo[q] = a;
    }
  }
}

// Lexicographic index of (x, y, z).
__device__ inline int64_t lex(const Geom& g, int x, int y, int z) {
  return (int64_t)x + (int64_t)g.L0 * ((int64_t)y + (int64_t)g.L1 * z);
}

// Population access of the fused step. pop_ld(p, q, V, g) is p[q V + g].
// For 4-byte populations the base of population q (the same for every lane)
// is pinned to SGPRs and the site is a 32-bit byte offset, so the compiler
// emits global_load v, v_off, s[base:base+1] with one offset register shared
// by the Q populations of a site; measured faster for FP32 storage. For
// 8-byte populations the plain 64-bit form measured faster (the FP64 MRT
// step is limited by its other registers, and the SGPR form changed its
// schedule for the worse), so it is kept. The 32-bit offset needs
// V sizeof(T) < 2^32, which simulation and batch creation check.
#define IDMX_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ uint64_t sgpr_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
template <class T>
__device__ __forceinline__ T pop_ld(const T* __restrict__ p, int q, int64_t V, int64_t g) {
  if constexpr (sizeof(T) == 4) {
    const IDMX_GLOBAL char* b = (const IDMX_GLOBAL char*)sgpr_u64((uint64_t)(p + (size_t)q * V));
    return *(const IDMX_GLOBAL T*)(b + (uint32_t)g * (uint32_t)sizeof(T));
  } else {
    return p[(size_t)q * V + g];
  }
}
template <class T>
__device__ __forceinline__ void pop_st_nt(T* __restrict__ p, int q, int64_t V, int64_t g, T v) {
  if constexpr (sizeof(T) == 4) {
    IDMX_GLOBAL char* b = (IDMX_GLOBAL char*)sgpr_u64((uint64_t)(p + (size_t)q * V));
    __builtin_nontemporal_store(v, (IDMX_GLOBAL T*)(b + (uint32_t)g * (uint32_t)sizeof(T)));
  } else {
    __builtin_nontemporal_store(v, p + (size_t)q * V + g);
  }
}

// Non-temporal store at a wave-uniform 64-bit base (SGPRs) plus a 32-bit
// per-lane byte offset: the compiler emits global_store v_off, v_data,
// s[base:base+1], so a site's Q stores share one offset register.
template <class T>
__device__ __forceinline__ void st_nt_sbase(uint64_t base, uint32_t off, T v) {
  IDMX_GLOBAL char* b = (IDMX_GLOBAL char*)sgpr_u64(base);
  __builtin_nontemporal_store(v, (IDMX_GLOBAL T*)(b + off));
}

// Push-stream one site's f* to its neighbours (periodic).
template <class S, class T>
__device__ inline void push_store(T* __restrict__ dst, const Geom& geo, int x,
                                  int y, int z, const T (&fo)[S::Q]) {
  // opaque site coordinates: the 19 target offsets are recomputed at every
  // call instead of being hoisted out of the plane loop into ~2Q VGPRs
  asm volatile("" : "+v"(x), "+v"(y));
#pragma unroll
  for (int q = 0; q < S::Q; q++) {
    const int xt = S::cx[q] ? wrap(x + S::cx[q], geo.L0) : x;
    const int yt = S::cy[q] ? wrap(y + S::cy[q], geo.L1) : y;
    const int zt = S::cz[q] ? wrap(z + S::cz[q], geo.L2) : z;
    pop_st_nt(dst, q, geo.V, lex(geo, xt, yt, zt), fo[q]);
  }
}

// Pull form of the populations: P_q(x) = f*_q(x) is the post-collision value
// kept at its source site, so that f_q(x, t+1) = P_q(x - c_q). A step in
// this form reads shifted (P_q at x - c_q: coalesced loads offset by one
// element cost nothing measurable) and stores every population at its own
// site (aligned), where the push form stores shifted, which costs ~17 % of
// a 19-stream copy. pull_src is the source of population q of the site at
// in-plane position (sx, sy) of marching plane mj.
template <class S>
__device__ __forceinline__ int64_t pull_src(const Geom& geo, int q, int sx, int sy, int mj) {
  if constexpr (S::D == 3) {
    const int xs = S::cx[q] ? wrap(sx - S::cx[q], geo.L0) : sx;
    const int ys = S::cy[q] ? wrap(sy - S::cy[q], geo.L1) : sy;
    const int zs = S::cz[q] ? wrap(mj - S::cz[q], geo.L2) : mj;
    return lex(geo, xs, ys, zs);
  } else {
    const int xs = S::cx[q] ? wrap(sx - S::cx[q], geo.L0) : sx;
    const int ys = S::cy[q] ? wrap(mj - S::cy[q], geo.L1) : mj;
    return (int64_t)xs + (int64_t)geo.L0 * ys;
  }
}

// Store one site's f* at the site itself (pull form).
template <class S, class T>
__device__ inline void local_store(T* __restrict__ dst, const Geom& geo, int x, int y, int z,
                                   const T (&fo)[S::Q]) {
  const int64_t g = lex(geo, x, y, z);
#pragma unroll
  for (int q = 0; q < S::Q; q++) pop_st_nt(dst, q, geo.V, g, fo[q]);
}

// Conversion between streamed populations and the pull form:
// dir = +1: dst_q(x) = src_q(x + c_q) (streamed -> pull form),
// dir = -1: dst_q(x) = src_q(x - c_q) (pull form -> streamed).
template <class S, class T>
__global__ void __launch_bounds__(256)
k_restream(const T* __restrict__ src, T* __restrict__ dst, Geom geo, int dir) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t L01 = (int64_t)geo.L0 * geo.L1;
    const int z = (int)(g / L01);
    const int64_t r = g - (int64_t)z * L01;
    const int y = (int)(r / geo.L0), x = (int)(r - (int64_t)y * geo.L0);
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const int xs = S::cx[q] ? wrap(x + dir * S::cx[q], geo.L0) : x;
      const int ys = S::cy[q] ? wrap(y + dir * S::cy[q], geo.L1) : y;
      const int zs = S::cz[q] ? wrap(z + dir * S::cz[q], geo.L2) : z;
      dst[(size_t)q * geo.V + g] = src[(size_t)q * geo.V + lex(geo, xs, ys, zs)];
    }
  }
}

// --------------------------------------------------------------------------
// Fused single-pass step.
//
// A block owns a TX x TY tile of the plane orthogonal to the slowest axis
// (z in 3D; in 2D the "plane" is an x-row of TX sites and TY = 1) and marches
// H planes along that axis. For every plane it reads ahead ("lead plane", R
// planes ahead of the one it collides) the populations of the tile plus an
// R-wide halo, sums them to n and stores psi(n) in an LDS ring of 2R+2
// planes. The tile's own populations stay in registers for R iterations
// until their plane is collided, so each population is read from HBM once
// per step except for the halo sites and the 2R extra planes per segment.
// One barrier per plane: the ring has one slot more than the stencil reach
// needs, so a plane's slot is rewritten only after every thread has passed
// the next barrier.
//
// Blocks are dispatched round-robin over the 8 XCDs; the tile order is
// remapped so that each XCD walks a contiguous range of tiles (neighbouring
// tiles share halo rows, which then hit the same L2).
//
// HIST selects where the tile's populations wait between their lead-plane
// read and their collision: in registers (true; one HBM read per value, but
// (R+1) Q values per thread held across the barrier), or nowhere (false; the
// collision re-reads them, R iterations later, from L2 / the infinity
// cache, which keeps the register count low enough for more waves per SIMD).
//
// PF (with HIST = false) issues the loads of the next lead plane right after
// the barrier, before the collision, so their latency overlaps the collision
// arithmetic and stores.
//
// PULL runs the step on the pull form (pull_src above): every population
// read is taken from its source site, and f* is stored at the site itself.
template <class S, class ES, class T, bool SHIFT, int COLL, bool WF, int TX, int TY,
          bool HIST, bool PF, bool PULL = false, bool P2 = false>
__device__ __forceinline__ void fused_body(const T* __restrict__ pop, T* __restrict__ pop_swap,
                                           const Geom& geo, const Phys<T>& P, int H, int ntx,
                                           int nty, int ntiles, double* __restrict__ n_out,
                                           double* __restrict__ psi_out,
                                           double* __restrict__ u_out) {
  static_assert(!(PF && HIST), "prefetch is for the re-read variant");
  static_assert(!(P2 && (HIST || PF)), "two planes per barrier is a re-read variant");
  constexpr int D = S::D, Q = S::Q, R = ES::R;
  constexpr int NT = TX * TY;
  constexpr int PX = TX + 2 * R;
  constexpr int PY = D == 3 ? TY + 2 * R : 1;
  // ring of psi planes: one slot more than the stencil reach needs (two
  // more with P2, which writes two planes before each barrier)
  constexpr int NS = P2 ? 2 * R + 4 : 2 * R + 2;
  constexpr int NH = PX * PY - NT;  // halo sites per plane
  static_assert(D == 3 || TY == 1, "2D tiles are rows");
  constexpr int NHT = (NH + NT - 1) / NT;  // halo sites per thread
  __shared__ T lds[NS][PY][PX];

  // XCD-aware remap of the (padded to a multiple of 8) grid.
  const int per = gridDim.x >> 3;
  const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (tile >= ntiles) return;

  int t = tile;
  const int txi = t % ntx;
  t /= ntx;
  const int tyi = D == 3 ? t % nty : 0;
  const int seg = D == 3 ? t / nty : t;
  const int x0 = txi * TX, y0 = tyi * TY;
  const int Lm = D == 3 ? geo.L2 : geo.L1;          // marching axis length
  const int64_t PS = D == 3 ? (int64_t)geo.L0 * geo.L1 : (int64_t)geo.L0;
  const int m0 = seg * H;
  const int Hs = min(H, Lm - m0);

  const int tid = threadIdx.x;
  const int ix = tid % TX, iy = tid / TX;
  const int x = x0 + ix, y = D == 3 ? y0 + iy : 0;
  const bool own_valid = x < geo.L0 && (D == 2 || y < geo.L1);
  const int xw = x % geo.L0, yw = D == 3 ? y % geo.L1 : 0;
  const int64_t own_off = (int64_t)xw + (int64_t)geo.L0 * yw;

  // This thread's halo sites: position in the haloed plane and in-plane
  // offset. Halo site h is enumerated as: R top rows, R bottom rows, then
  // the R left / R right columns of the TY middle rows (2D: R left, R right).
  int hxs[NHT], hys[NHT];
  int64_t hoff[NHT];
#pragma unroll
  for (int k = 0; k < NHT; k++) {
    const int h = tid + k * NT;
    int hx = 0, hy = 0;
    if (D == 2) {
      hx = h < R ? h : TX + h;
    } else {
      const int top = R * PX;
      if (h < top) {
        hy = h / PX; hx = h % PX;
      } else if (h < 2 * top) {
        hy = R + TY + (h - top) / PX; hx = (h - top) % PX;
      } else {
        const int kk = h - 2 * top;
        hy = R + kk / (2 * R);
        const int sd = kk % (2 * R);
        hx = sd < R ? sd : TX + sd;
      }
    }
    hxs[k] = hx;
    hys[k] = hy;
    const int gx = ((x0 - R + hx) % geo.L0 + geo.L0) % geo.L0;
    const int gy = D == 3 ? ((y0 - R + hy) % geo.L1 + geo.L1) % geo.L1 : 0;
    hoff[k] = PULL ? (int64_t)gx + ((int64_t)gy << 32)  // (x, y) of the halo site
                   : (int64_t)gx + (int64_t)geo.L0 * gy;
  }

  constexpr int HR = HIST ? R + 1 : 1;
  T fh[HR][Q];  // own populations of the last R+1 lead planes (HIST)
  T dnh[HR];    // their sums
  T fn[PF ? Q : 1];  // prefetched own populations of the next lead plane
  // population q of this thread's own site in marching plane mj. In the
  // pull form the source offset is rebuilt from the site's in-plane offset
  // and its four edge flags for every plane (a few adds and selects), so
  // that the compiler does not keep 19 hoisted offsets live in VGPRs.
  const bool ex0 = xw == 0, exL = xw == geo.L0 - 1;
  const bool ey0 = D == 3 && yw == 0, eyL = D == 3 && yw == geo.L1 - 1;
  auto own_ld = [&](int q, int mj) -> T {
    if constexpr (PULL) {
      int64_t o = own_off;
      asm volatile("" : "+v"(o));  // opaque: recomputed per plane, not hoisted
      o -= S::cx[q] + (D == 3 ? (int64_t)geo.L0 * S::cy[q] : 0);
      if (S::cx[q] > 0 && ex0) o += geo.L0;
      if (S::cx[q] < 0 && exL) o -= geo.L0;
      if (D == 3 && S::cy[q] > 0 && ey0) o += PS;
      if (D == 3 && S::cy[q] < 0 && eyL) o -= PS;
      const int cm = D == 3 ? S::cz[q] : S::cy[q];
      const int ms = cm ? wrap(mj - cm, Lm) : mj;
      return pop_ld(pop, q, geo.V, (int64_t)ms * PS + o);
    } else {
      return pop_ld(pop, q, geo.V, (int64_t)mj * PS + own_off);
    }
  };
  if (PF) {
    const int mp = ((m0 - R) % Lm + Lm) % Lm;
#pragma unroll
    for (int q = 0; q < Q; q++) fn[q] = own_ld(q, mp);
  }

  if constexpr (P2) {
    // Two lead planes per barrier: the loads of both planes are in flight
    // together and the per-plane barrier cost is halved. Same arithmetic.
    const int J = Hs + 2 * R;
    auto lead = [&](int j) {
      const int mj = ((m0 - R + j) % Lm + Lm) % Lm;
      const int64_t base = (int64_t)mj * PS;
      const int slot = j % NS;
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++) s += own_ld(q, mj);
      lds[slot][D == 3 ? R + iy : 0][R + ix] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
#pragma unroll
      for (int k = 0; k < NHT; k++) {
        if (tid + k * NT < NH) {
          T sh = 0;
#pragma unroll
          for (int q = 0; q < Q; q++) {
            if constexpr (PULL) {
              int64_t hp = hoff[k];
              asm volatile("" : "+v"(hp));
              sh += pop_ld(pop, q, geo.V,
                           pull_src<S>(geo, q, (int)(hp & 0xffffffff), (int)(hp >> 32), mj));
            } else {
              sh += pop_ld(pop, q, geo.V, base + hoff[k]);
            }
          }
          lds[slot][hys[k]][hxs[k]] = psi_of(P.psi_kind, SHIFT ? P.rho0 + sh : sh);
        }
      }
    };
    auto coll = [&](int j) {
      const int mc = m0 + j - 2 * R;
      T f[Q];
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++) f[q] = own_ld(q, mc);
#pragma unroll
      for (int q = 0; q < Q; q++) s += f[q];
      const int cy = D == 3 ? R + iy : 0;
      const int cx = R + ix;
      T gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int e = 0; e < ES::QE; e++) {
        const int em = D == 3 ? ES::ez[e] : ES::ey[e];
        const int ey = D == 3 ? ES::ey[e] : 0;
        const T v = lds[(j - R + em) % NS][cy + ey][cx + ES::ex[e]];
        const T w = T(ES::w[e]);
        if (ES::ex[e] != 0) gx += T(ES::ex[e]) * w * v;
        if (ES::ey[e] != 0) gy += T(ES::ey[e]) * w * v;
        if (ES::ez[e] != 0) gz += T(ES::ez[e]) * w * v;
      }
      const T lpsi = lds[(j - R) % NS][cy][cx];
      T fo[Q], u[3];
      collide<S, T, SHIFT, COLL>(f, s, gx, gy, gz, lpsi, P, fo, u);
      const int xc = x, yc = D == 3 ? y : mc, zc = D == 3 ? mc : 0;
      if constexpr (PULL) local_store<S, T>(pop_swap, geo, xc, yc, zc, fo);
      else push_store<S, T>(pop_swap, geo, xc, yc, zc, fo);
      if (WF) {
        const int64_t g = lex(geo, xc, yc, zc);
        n_out[g] = (double)(SHIFT ? P.rho0 + s : s);
        psi_out[g] = (double)lpsi;
        u_out[g] = (double)u[0];
        u_out[g + geo.V] = (double)u[1];
        if (D == 3) u_out[g + 2 * geo.V] = (double)u[2];
      }
    };
    for (int j = 0; j < J; j += 2) {
      lead(j);
      if (j + 1 < J) lead(j + 1);
      __syncthreads();
      if (own_valid && j >= 2 * R) coll(j);
      if (own_valid && j + 1 < J && j + 1 >= 2 * R) coll(j + 1);
    }
    return;
  }

  for (int j = 0; j < Hs + 2 * R; j++) {
    const int mj = ((m0 - R + j) % Lm + Lm) % Lm;
    const int64_t base = (int64_t)mj * PS;
    const int slot = j % NS;
    {
      T s = 0;
      if (PF) {
#pragma unroll
        for (int q = 0; q < Q; q++) fh[HR - 1][q] = fn[q];
      } else {
#pragma unroll
        for (int q = 0; q < Q; q++) fh[HR - 1][q] = own_ld(q, mj);
      }
#pragma unroll
      for (int q = 0; q < Q; q++) s += fh[HR - 1][q];
      dnh[HR - 1] = s;
      lds[slot][D == 3 ? R + iy : 0][R + ix] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
    }
#pragma unroll
    for (int k = 0; k < NHT; k++) {
      if (tid + k * NT < NH) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
          if constexpr (PULL) {
            int64_t hp = hoff[k];
            asm volatile("" : "+v"(hp));  // opaque: recomputed per plane, not hoisted
            s += pop_ld(pop, q, geo.V,
                        pull_src<S>(geo, q, (int)(hp & 0xffffffff), (int)(hp >> 32), mj));
          }
          else
            s += pop_ld(pop, q, geo.V, base + hoff[k]);
        }
        lds[slot][hys[k]][hxs[k]] = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
      }
    }
    __syncthreads();
    if (PF && j + 1 < Hs + 2 * R) {
      const int mn = ((m0 - R + j + 1) % Lm + Lm) % Lm;
#pragma unroll
      for (int q = 0; q < Q; q++) fn[q] = own_ld(q, mn);
    }
    if (j >= 2 * R && own_valid) {
      // plane with lead index j - R
      const int mc = m0 + j - 2 * R;
      if (!HIST) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) fh[0][q] = own_ld(q, mc);
#pragma unroll
        for (int q = 0; q < Q; q++) s += fh[0][q];
        dnh[0] = s;
      }
      int sl[2 * R + 1];
#pragma unroll
      for (int k = 0; k <= 2 * R; k++) sl[k] = (j - 2 * R + k) % NS;
      const int cy = D == 3 ? R + iy : 0;
      const int cx = R + ix;
      T gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int e = 0; e < ES::QE; e++) {
        const int em = D == 3 ? ES::ez[e] : ES::ey[e];
        const int ey = D == 3 ? ES::ey[e] : 0;
        const T v = lds[sl[R + em]][cy + ey][cx + ES::ex[e]];
        const T w = T(ES::w[e]);
        if (ES::ex[e] != 0) gx += T(ES::ex[e]) * w * v;
        if (ES::ey[e] != 0) gy += T(ES::ey[e]) * w * v;
        if (ES::ez[e] != 0) gz += T(ES::ez[e]) * w * v;
      }
      const T lpsi = lds[sl[R]][cy][cx];
      T fo[Q], u[3];
      collide<S, T, SHIFT, COLL>(fh[0], dnh[0], gx, gy, gz, lpsi, P, fo, u);
      const int xc = x, yc = D == 3 ? y : mc, zc = D == 3 ? mc : 0;
      if constexpr (PULL) local_store<S, T>(pop_swap, geo, xc, yc, zc, fo);
      else push_store<S, T>(pop_swap, geo, xc, yc, zc, fo);
      if (WF) {
        const int64_t g = lex(geo, xc, yc, zc);
        const T dn = dnh[0];
        n_out[g] = (double)(SHIFT ? P.rho0 + dn : dn);
        psi_out[g] = (double)lpsi;
        u_out[g] = (double)u[0];
        u_out[g + geo.V] = (double)u[1];
        if (D == 3) u_out[g + 2 * geo.V] = (double)u[2];
      }
    }
    if (HIST) {
#pragma unroll
      for (int k = 0; k < HR - 1; k++) {
        dnh[k] = dnh[k + 1];
#pragma unroll
        for (int q = 0; q < Q; q++) fh[k][q] = fh[k + 1][q];
      }
    }
  }
}

template <class S, class ES, class T, bool SHIFT, int COLL, bool WF, int TX, int TY,
          bool HIST = true, bool PF = false, bool PULL = false, bool P2 = false>
__global__ void __launch_bounds__(TX * TY)
k_fused(const T* __restrict__ pop, T* __restrict__ pop_swap, Geom geo, Phys<T> P, int H,
        int ntx, int nty, int ntiles, double* __restrict__ n_out, double* __restrict__ psi_out,
        double* __restrict__ u_out) {
  fused_body<S, ES, T, SHIFT, COLL, WF, TX, TY, HIST, PF, PULL, P2>(
      pop, pop_swap, geo, P, H, ntx, nty, ntiles, n_out, psi_out, u_out);
}

// Same kernel compiled for at least 3 waves per SIMD (the register budget
// drops to 168 per lane).
template <class S, class ES, class T, bool SHIFT, int COLL, bool WF, int TX, int TY,
          bool HIST = true, bool PF = false>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(3)))
k_fused_o3(const T* __restrict__ pop, T* __restrict__ pop_swap, Geom geo, Phys<T> P, int H,
           int ntx, int nty, int ntiles, double* __restrict__ n_out,
           double* __restrict__ psi_out, double* __restrict__ u_out) {
  fused_body<S, ES, T, SHIFT, COLL, WF, TX, TY, HIST, PF>(pop, pop_swap, geo, P, H, ntx, nty,
                                                          ntiles, n_out, psi_out, u_out);
}

// --------------------------------------------------------------------------
// Fused step with the lead plane staged in LDS by DMA (push form).
//
// Same march as fused_body (tile of TX x TY sites, H planes, psi ring in
// LDS, the tile's own populations held in registers for R planes), but the
// lead plane's populations -- tile and R-wide halo, all Q of them -- are
// copied from HBM into LDS with global_load_lds_dword (the load unit writes
// LDS directly, no VGPRs in flight), and the copy of plane j+1 is issued
// before the collision of plane j-R, so it overlaps that arithmetic and its
// stores. Per plane:
//   wait for the copy of plane j; barrier;
//   own populations LDS -> registers, psi(n) of tile and halo -> ring;
//   barrier (the staging buffer is free again);
//   issue the copy of plane j+1; collide plane j-R and push its f*.
//
// Waits: at the top of an iteration every wave waits for all of its vector
// memory operations (vmcnt(0)): the copy and the push stores issued behind
// it. This makes no assumption about the order in which loads and stores
// complete; waiting only for the copy (vmcnt(Q), counting on in-order
// completion) measured no faster, since the stores were issued a collision
// earlier. The barriers wait for LDS only (lgkmcnt).
//
// Staging layout: population q of the haloed plane occupies a block of NK
// chunks of 64 dwords (one DMA instruction each); dword d of the block is
// dword d % EPD of haloed site d / EPD (EPD = sizeof(T) / 4), sites in
// row-major order of the (TX + 2R) x (TY + 2R) haloed tile. Chunk k of every
// population is fetched by wave k % NW with the same per-lane source offset,
// precomputed once per block (the periodic wrap of the halo included).
// Copies of 16 bytes per lane (global_load_lds_dwordx4, with rows widened
// to 16-byte groups) measured 15-20 % slower on this step.
template <class S, class ES, class T, int TX, int TY, int COLL, bool WALLS = false>
struct DmaShape {
  static constexpr int D = S::D, Q = S::Q, R = ES::R;
  // with walls two more planes are staged per lead plane, each laid out as
  // a population: the psi the caller stored (used at solid sites) and the
  // per-site wall masks (64-bit)
  static constexpr int NQ = Q + (WALLS ? 2 : 0);
  static constexpr int NT = TX * TY;
  static constexpr int NW = NT / 64;
  static constexpr int PX = TX + 2 * R;
  static constexpr int PY = D == 3 ? TY + 2 * R : 1;
  static constexpr int NP = PX * PY;                 // haloed sites per plane
  static constexpr int EPD = (int)sizeof(T) / 4;     // dwords per population value
  static constexpr int DW = NP * EPD;                // dwords per population
  static constexpr int NK = (DW + 63) / 64;          // DMA chunks per population
  static constexpr int KW = (NK + NW - 1) / NW;      // chunks per wave and population
  static constexpr int NS = 2 * R + 1;               // psi ring slots
  static constexpr int QB = NK * 256;                // bytes per population block
  static constexpr int NB = NBasis<D>::value;
  // collision tables copied to LDS: MRT rates (diagonal S) or A, B, G (dense)
  static constexpr int NTAB = COLL == 2 ? Q : COLL == 1 ? Q * Q + 2 * NB * Q : 1;
  static constexpr int LDS = NQ * QB + NS * NP * (int)sizeof(T) + NTAB * (int)sizeof(T);
  static constexpr bool fits =
      LDS <= 160 * 1024 && NT % 64 == 0 && (!WALLS || sizeof(T) == sizeof(double));
};

// Wall mask of a site for the fused step with walls: bit q (q >= 1) when the
// neighbour x + c_q is solid (flag 0: the population pushed there bounces
// back), bit 63 when the site itself is not fluid (flag != 1: not updated,
// its psi is the one the caller stored).
constexpr int kWallSelf = 63;

__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One dword per lane from global memory (SGPR base + 32-bit VGPR offset)
// into LDS at lds + 4 * lane. Written as inline assembly so that the
// compiler does not track it: it would otherwise wait for vmcnt(0) before
// every LDS read after it, stores included. The caller waits for it
// explicitly. m0 (the LDS base of the transfer) is saved and restored.
__device__ __forceinline__ void dma_dword(uint64_t base, uint32_t voff, uint32_t lds) {
  uint32_t keep;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %1\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dword %2, %3\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(keep)
      : "s"(lds), "v"(voff), "s"(base)
      : "memory");
}

// WALLS: the step of MainLoopWalls (fluid sites only, half-way bounce-back
// at solid sites, the caller's psi at non-fluid sites; K_ForceCollide-
// StreamWallsSCMPMeta, idpy/LBM/MultiPhaseKernelsMeta.py:524-716) with a
// body force (bx, by, bz) added to the Shan-Chen force: psi_w is the psi
// array (its non-fluid sites are staged with the populations), wmask the
// per-site masks (kWallSelf). FP64 only.
template <class S, class ES, class T, bool SHIFT, int COLL, int TX, int TY, bool WALLS = false,
          bool XSH = false>
__global__ void __launch_bounds__(TX * TY)
k_fused_dma(const T* __restrict__ pop, T* __restrict__ pop_swap, Geom geo, Phys<T> P, int H,
            int ntx, int nty, int ntiles, const double* __restrict__ psi_w = nullptr,
            const uint64_t* __restrict__ wmask = nullptr, T bx = T(0), T by = T(0),
            T bz = T(0), int xsh = 0) {
  using SH = DmaShape<S, ES, T, TX, TY, COLL, WALLS>;
  constexpr int D = S::D, Q = S::Q, R = ES::R;
  constexpr int NT = SH::NT, PX = SH::PX, PY = SH::PY, NP = SH::NP, NS = SH::NS;
  constexpr int NK = SH::NK, KW = SH::KW, NW = SH::NW, EPD = SH::EPD, DW = SH::DW;
  static_assert(SH::fits, "staging buffer exceeds the LDS of a CU");
  static_assert(Q <= 63, "vmcnt holds at most 63 outstanding operations");
  __shared__ __attribute__((aligned(16))) char stage[SH::NQ * SH::QB];
  __shared__ T ring[NS][PY][PX];
  __shared__ T tab[SH::NTAB];

  const int per = gridDim.x >> 3;
  const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  int t = tile;
  const int txi = t % ntx;
  t /= ntx;
  const int tyi = D == 3 ? t % nty : 0;
  const int seg = D == 3 ? t / nty : t;
  const int x0 = txi * TX, y0 = tyi * TY;
  const int Lm = D == 3 ? geo.L2 : geo.L1;
  const int64_t PS = D == 3 ? (int64_t)geo.L0 * geo.L1 : (int64_t)geo.L0;
  const int m0 = seg * H;
  const int Hs = min(H, Lm - m0);
  const int J = Hs + 2 * R;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);

  // collision tables in LDS: read there with LDS loads, not with vector
  // memory loads, whose completion would have to wait for the plane copy
  Phys<T> PL = P;
  static_assert(COLL >= 0 && COLL <= 4 && !(WALLS && COLL > 2),
                "the LDS-staged step runs BGK and MRT with the Guo forcing, BGK with the "
                "Shan-Chen forcing and BGK with the isotropy filter");
  if constexpr (COLL == 1 || COLL == 2) {
    for (int i = tid; i < SH::NTAB; i += NT) tab[i] = P.mrtA[i];
    PL.mrtA = tab;
    PL.mrtB = COLL == 1 ? tab + Q * Q : nullptr;
    PL.mrtG = COLL == 1 ? tab + Q * Q + SH::NB * Q : nullptr;
  }
  const int ix = tid % TX, iy = tid / TX;
  const int x = x0 + ix, y = D == 3 ? y0 + iy : 0;
  const bool own_valid = x < geo.L0 && (D == 2 || y < geo.L1);

  // Per-lane source byte offsets (within a plane of one population) of the
  // chunks this wave fetches, and whether the lane's dword exists.
  uint32_t soff[KW];
  // XSH: the x-shifted storage, in which population q of site x is kept at
  // x - c_qx. Stores into it cover the tile's own, line-aligned x range for
  // every population (streamed storage: x + c_qx, one site off the lines for
  // the populations with c_qx != 0); copies from it read the populations
  // with c_qx = +1 / -1 from x - 1 / x + 1. xsh bit 0: the step reads the
  // shifted storage, bit 1: it writes it (a run converts on its first and
  // last step, with no pass of its own). With walls a population that
  // bounces back goes to its own site as the opposite population, which
  // that storage keeps at x + c_qx; the sites that are not fluid are never
  // written, in either storage.
  const bool rsh = XSH && (xsh & 1), wsh = XSH && (xsh & 2);
  uint32_t soffp[XSH ? KW : 1], soffm[XSH ? KW : 1];
  bool sval[KW];
#pragma unroll
  for (int kk = 0; kk < KW; kk++) {
    const int k = wv + kk * NW;
    const int d = k * 64 + lane;
    sval[kk] = k < NK && d < DW;
    const int site = sval[kk] ? d / EPD : 0;
    const int hx = site % PX, hy = site / PX;
    const int gx = ((x0 - R + hx) % geo.L0 + geo.L0) % geo.L0;
    const int gy = D == 3 ? ((y0 - R + hy) % geo.L1 + geo.L1) % geo.L1 : 0;
    soff[kk] = ((uint32_t)gx + (uint32_t)geo.L0 * (uint32_t)gy) * (uint32_t)sizeof(T) +
               (uint32_t)(d % EPD) * 4u;
    if constexpr (XSH) {
      // the streamed storage's offsets where this step reads it
      const uint32_t gxp = rsh ? (uint32_t)wrap(gx - 1, geo.L0) : (uint32_t)gx;
      const uint32_t gxm = rsh ? (uint32_t)wrap(gx + 1, geo.L0) : (uint32_t)gx;
      soffp[kk] = (gxp + (uint32_t)geo.L0 * (uint32_t)gy) * (uint32_t)sizeof(T) +
                  (uint32_t)(d % EPD) * 4u;
      soffm[kk] = (gxm + (uint32_t)geo.L0 * (uint32_t)gy) * (uint32_t)sizeof(T) +
                  (uint32_t)(d % EPD) * 4u;
    }
  }
  const uint64_t qstride = (uint64_t)geo.V * sizeof(T);
  const uint32_t stage_lds = __builtin_amdgcn_readfirstlane(
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) char*)stage);
  auto issue_copy = [&](int j) {
    const int mj = ((m0 - R + j) % Lm + Lm) % Lm;
    // base of population q in plane mj, stepped by one population per q and
    // kept opaque so that the Q bases are not all precomputed into SGPRs
    uint64_t base = sgpr_u64((uint64_t)(pop + (size_t)mj * (size_t)PS));
#pragma unroll
    for (int q = 0; q < Q; q++) {
      if (q) base += qstride;
      asm volatile("" : "+s"(base));
#pragma unroll
      for (int kk = 0; kk < KW; kk++) {
        const int k = wv + kk * NW;
        if (k < NK && sval[kk])
          dma_dword(base,
                    !XSH || S::cx[q] == 0 ? soff[kk]
                                          : S::cx[q] > 0 ? soffp[kk % (XSH ? KW : 1)]
                                                         : soffm[kk % (XSH ? KW : 1)],
                    stage_lds + (uint32_t)(q * SH::QB + k * 256));
      }
    }
    if constexpr (WALLS) {
      // the caller's psi plane and the wall-mask plane, 8 bytes per site
      // like an FP64 population
      const uint64_t bp = sgpr_u64((uint64_t)(psi_w + (size_t)mj * (size_t)PS));
      const uint64_t bm = sgpr_u64((uint64_t)(wmask + (size_t)mj * (size_t)PS));
#pragma unroll
      for (int kk = 0; kk < KW; kk++) {
        const int k = wv + kk * NW;
        if (k < NK && sval[kk]) {
          dma_dword(bp, soff[kk], stage_lds + (uint32_t)(Q * SH::QB + k * 256));
          dma_dword(bm, soff[kk], stage_lds + (uint32_t)((Q + 1) * SH::QB + k * 256));
        }
      }
    }
  };
  // staged psi and mask of a haloed site (walls only)
  auto st_psiw = [&](int site) -> T {
    return *(const T*)(stage + Q * SH::QB + site * (int)sizeof(T));
  };
  auto st_mask = [&](int site) -> uint64_t {
    return *(const uint64_t*)(stage + (Q + 1) * SH::QB + site * 8);
  };
  auto st_val = [&](int q, int site) -> T {
    return *(const T*)(stage + q * SH::QB + site * (int)sizeof(T));
  };

  constexpr int HR = R + 1;
  T fh[HR][Q];
  T dnh[HR];
  uint64_t mh[HR];  // walls: the own site's mask, with the populations' history
  const int own_site = (D == 3 ? (iy + R) * PX : 0) + ix + R;

  // Push targets: byte offsets within a marching plane of the site and its
  // in-plane neighbours (periodic), computed once; the target plane of
  // population q is wave-uniform and goes into the scalar base.
  const int xm = wrap(x - 1, geo.L0), xp = wrap(x + 1, geo.L0);
  const int ym = D == 3 ? wrap(y - 1, geo.L1) : 0, yp = D == 3 ? wrap(y + 1, geo.L1) : 0;
  uint32_t toff[3][3];  // [dy + 1][dx + 1]
#pragma unroll
  for (int b = 0; b < 3; b++) {
    const int yy = b == 0 ? ym : b == 1 ? y : yp;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int xx = a == 0 ? xm : a == 1 ? x : xp;
      // where this step writes the x-shifted storage every population of
      // the site goes to its own x (with walls the streamed offsets are
      // kept: the bounce-back stores use them)
      toff[b][a] = ((uint32_t)(wsh && !WALLS ? x : xx) +
                    (uint32_t)geo.L0 * (uint32_t)(D == 3 ? yy : 0)) *
                   (uint32_t)sizeof(T);
    }
  }
  const uint64_t qbytes = (uint64_t)geo.V * sizeof(T);
  // f* of the last collided plane, stored after the next copy is issued so
  // that the stores and the copy's loads are in flight together
  T fpend[Q];
  int pend_m = -1;
  uint64_t pend_mask = 0;
  auto flush = [&]() {
    if constexpr (WALLS) {
      // fluid sites only; a population aimed at a solid site goes back to
      // its own site as the opposite population
      if (pend_m >= 0 && own_valid && !((pend_mask >> kWallSelf) & 1)) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int cm = D == 3 ? S::cz[q] : S::cy[q];
          const int mt = cm ? wrap(pend_m + cm, Lm) : pend_m;
          const int dy = D == 3 ? S::cy[q] : 0;
          const uint64_t a_push = (uint64_t)(pop_swap) + (uint64_t)q * qbytes +
                                  (uint64_t)mt * (uint64_t)PS * sizeof(T) +
                                  toff[dy + 1][wsh ? 1 : S::cx[q] + 1];
          const uint64_t a_back = (uint64_t)(pop_swap) + (uint64_t)opposite<S>(q) * qbytes +
                                  (uint64_t)pend_m * (uint64_t)PS * sizeof(T) +
                                  toff[1][wsh ? S::cx[q] + 1 : 1];
          const bool back = q != 0 && ((pend_mask >> q) & 1);
          __builtin_nontemporal_store(fpend[q], (IDMX_GLOBAL T*)(back ? a_back : a_push));
        }
      }
    } else if (pend_m >= 0 && own_valid) {
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int cm = D == 3 ? S::cz[q] : S::cy[q];
        const int mt = cm ? wrap(pend_m + cm, Lm) : pend_m;
        const int dy = D == 3 ? S::cy[q] : 0;
        uint64_t base = (uint64_t)(pop_swap) + (uint64_t)q * qbytes +
                        (uint64_t)mt * (uint64_t)PS * sizeof(T);
        st_nt_sbase(base, toff[dy + 1][S::cx[q] + 1], fpend[q]);
      }
    }
    pend_m = -1;
  };

  issue_copy(0);
  for (int j = 0; j < J; j++) {
    // raised wave priority from here until the next copy is issued: this
    // stretch is on the critical path of the block, while the other block
    // on the CU may be colliding (measured 1-5 % faster, steadier)
    asm volatile("s_setprio 2" ::: "memory");
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    lds_barrier();
    const int slot = j % NS;
    {
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        fh[HR - 1][q] = st_val(q, own_site);
        s += fh[HR - 1][q];
      }
      dnh[HR - 1] = s;
      T ps = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
      if constexpr (WALLS) {
        mh[HR - 1] = st_mask(own_site);
        if ((mh[HR - 1] >> kWallSelf) & 1) ps = st_psiw(own_site);
      }
      ring[slot][D == 3 ? R + iy : 0][R + ix] = ps;
    }
    // halo sites: haloed index h of the ring enumerated over the NP - NT
    // sites outside the tile, as in fused_body
    for (int h = tid; h < NP - NT; h += NT) {
      int hx = 0, hy = 0;
      if (D == 2) {
        hx = h < R ? h : TX + h;
      } else {
        const int top = R * PX;
        if (h < top) {
          hy = h / PX; hx = h % PX;
        } else if (h < 2 * top) {
          hy = R + TY + (h - top) / PX; hx = (h - top) % PX;
        } else {
          const int kk = h - 2 * top;
          hy = R + kk / (2 * R);
          const int sd = kk % (2 * R);
          hx = sd < R ? sd : TX + sd;
        }
      }
      const int site = hy * PX + hx;
      T s = 0;
#pragma unroll
      for (int q = 0; q < Q; q++) s += st_val(q, site);
      T ps = psi_of(P.psi_kind, SHIFT ? P.rho0 + s : s);
      if constexpr (WALLS)
        if ((st_mask(site) >> kWallSelf) & 1) ps = st_psiw(site);
      ring[slot][hy][hx] = ps;
    }
    lds_barrier();
    if (j + 1 < J) issue_copy(j + 1);
    asm volatile("s_setprio 0" ::: "memory");
    // the previous plane's stores, issued behind the copy
    flush();
    if (j >= 2 * R) {
      if (own_valid) {
        const int mc = m0 + j - 2 * R;
        const int cy = D == 3 ? R + iy : 0;
        const int cx = R + ix;
        T gx = 0, gy = 0, gz = 0;
#pragma unroll
        for (int e = 0; e < ES::QE; e++) {
          const int em = D == 3 ? ES::ez[e] : ES::ey[e];
          const int ey = D == 3 ? ES::ey[e] : 0;
          const T v = ring[(j - R + em) % NS][cy + ey][cx + ES::ex[e]];
          const T w = T(ES::w[e]);
          if (ES::ex[e] != 0) gx += T(ES::ex[e]) * w * v;
          if (ES::ey[e] != 0) gy += T(ES::ey[e]) * w * v;
          if (ES::ez[e] != 0) gz += T(ES::ez[e]) * w * v;
        }
        const T lpsi = ring[(j - R) % NS][cy][cx];
        T u[3];
        collide<S, T, SHIFT, COLL, WALLS>(fh[0], dnh[0], gx, gy, gz, lpsi, PL, fpend, u, bx, by,
                                          bz);
        if constexpr (WALLS) pend_mask = mh[0];
      }
      pend_m = m0 + j - 2 * R;
    }
#pragma unroll
    for (int k = 0; k < HR - 1; k++) {
      dnh[k] = dnh[k + 1];
      if constexpr (WALLS) mh[k] = mh[k + 1];
#pragma unroll
      for (int q = 0; q < Q; q++) fh[k][q] = fh[k + 1][q];
    }
  }
  flush();
}

// --------------------------------------------------------------------------
// Two-pass kernels with the reference's structure.

// Walls (MainLoopWalls, idpy/LBM/MultiPhase.py:1226-1309): a flag per site,
// 1 = fluid, 0 = solid. Only fluid sites are updated; a population pushed
// towards a solid site returns to its source site in the opposite direction
// (half-way bounce-back, idpy/LBM/Collisions.py:835-850); solid sites keep
// the psi the caller stored there (GetWalls, MultiPhase.py:157-161), which
// enters the force of their fluid neighbours. walls == nullptr: no walls.


// Push-stream with half-way bounce-back at solid sites.
template <class S, class T>
__device__ inline void push_store_walls(T* __restrict__ dst, const Geom& geo,
                                        const uint8_t* __restrict__ walls, int64_t g, int x,
                                        int y, int z, const T (&fo)[S::Q]) {
#pragma unroll
  for (int q = 0; q < S::Q; q++) {
    const int xt = S::cx[q] ? wrap(x + S::cx[q], geo.L0) : x;
    const int yt = S::cy[q] ? wrap(y + S::cy[q], geo.L1) : y;
    const int zt = S::cz[q] ? wrap(z + S::cz[q], geo.L2) : z;
    int64_t gt = lex(geo, xt, yt, zt);
    int qt = q;
    if (q != 0 && walls[gt] == 0) {
      gt = g;
      qt = opposite<S>(q);
    }
    dst[(size_t)qt * geo.V + gt] = fo[q];
  }
}

// K_ComputeDensityPsiMeta: n, psi from pop (MultiPhaseKernelsMeta.py:107-167;
// with walls K_ComputeDensityPsiWallsMeta, :42-104: fluid sites only).
// PROG: psi from the compiled psi_code program (a separate instantiation, so
// that the closed forms keep their registers and occupancy).
template <class S, class T, bool SHIFT, bool PROG = false>
__global__ void __launch_bounds__(256)
k_density_psi(const T* __restrict__ pop, Geom geo, Phys<T> P,
              double* __restrict__ n, double* __restrict__ psi,
              const uint8_t* __restrict__ walls) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (walls && walls[g] != 1) continue;
    T s = 0;
#pragma unroll
    for (int q = 0; q < S::Q; q++) s += pop[(size_t)q * geo.V + g];
    const T ln = SHIFT ? P.rho0 + s : s;
    n[g] = (double)ln;
    if (psi) psi[g] = PROG ? psi_prog(P.prog, (double)ln) : (double)psi_of(P.psi_kind, ln);
  }
}

// Gradient of psi at (x, y, z) gathered from a global psi array.
template <class S, class ES, class T>
__device__ inline void grad_global(const double* __restrict__ psi, const Geom& geo,
                                   int x, int y, int z, T& gx, T& gy, T& gz) {
  gx = gy = gz = 0;
#pragma unroll
  for (int e = 0; e < ES::QE; e++) {
    const int xn = ES::ex[e] ? wrap(x + ES::ex[e], geo.L0) : x;
    const int yn = ES::ey[e] ? wrap(y + ES::ey[e], geo.L1) : y;
    const int zn = ES::ez[e] ? wrap(z + ES::ez[e], geo.L2) : z;
    const T v = (T)psi[lex(geo, xn, yn, zn)];
    const T w = T(ES::w[e]);
    if (ES::ex[e] != 0) gx += T(ES::ex[e]) * w * v;
    if (ES::ey[e] != 0) gy += T(ES::ey[e]) * w * v;
    if (ES::ez[e] != 0) gz += T(ES::ez[e]) * w * v;
  }
}

__device__ inline void site_xyz(const Geom& geo, int64_t g, int& x, int& y, int& z) {
  x = (int)(g % geo.L0);
  const int64_t r = g / geo.L0;
  y = (int)(r % geo.L1);
  z = (int)(r / geo.L1);
}

// The per-site wall masks of the fused step with walls (kWallSelf), from the
// flags: recomputed at every run, so flags changed between runs are seen.
template <class S>
__global__ void __launch_bounds__(256)
k_wall_mask(const uint8_t* __restrict__ walls, Geom geo, uint64_t* __restrict__ mask) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    site_xyz(geo, g, x, y, z);
    uint64_t m = walls[g] != 1 ? (uint64_t(1) << kWallSelf) : uint64_t(0);
#pragma unroll
    for (int q = 1; q < S::Q; q++) {
      const int xt = S::cx[q] ? wrap(x + S::cx[q], geo.L0) : x;
      const int yt = S::cy[q] ? wrap(y + S::cy[q], geo.L1) : y;
      const int zt = S::cz[q] ? wrap(z + S::cz[q], geo.L2) : z;
      if (walls[lex(geo, xt, yt, zt)] == 0) m |= uint64_t(1) << q;
    }
    mask[g] = m;
  }
}

// K_ForceCollideStreamSCMPMeta / _MRTMeta (MultiPhaseKernelsMeta.py:169-345,
// 891-1075): force from the psi array, n from the n array, push stream.
// With walls, K_ForceCollideStreamWallsSCMPMeta (:524-716): fluid sites only,
// bounce-back at solid sites. (bx, by, bz) is a body force added to the
// Shan-Chen force (zero unless set).
template <class S, class ES, class T, bool SHIFT, int COLL, bool WF>
__global__ void __launch_bounds__(256)
k_fcs(const T* __restrict__ pop, T* __restrict__ pop_swap,
      const double* __restrict__ n_in, const double* __restrict__ psi_in,
      Geom geo, Phys<T> P, double* __restrict__ u_out, const uint8_t* __restrict__ walls,
      double bx, double by, double bz) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (walls && walls[g] != 1) continue;
    int x, y, z;
    site_xyz(geo, g, x, y, z);
    T gx, gy, gz;
    grad_global<S, ES, T>(psi_in, geo, x, y, z, gx, gy, gz);
    T f[S::Q];
#pragma unroll
    for (int q = 0; q < S::Q; q++) f[q] = pop[(size_t)q * geo.V + g];
    const T ln = (T)n_in[g];
    const T dn = SHIFT ? ln - P.rho0 : ln;
    T fo[S::Q], u[3];
    collide<S, T, SHIFT, COLL, true>(f, dn, gx, gy, gz, (T)psi_in[g], P, fo, u, (T)bx, (T)by,
                                     (T)bz);
    if (walls) push_store_walls<S, T>(pop_swap, geo, walls, g, x, y, z, fo);
    else push_store<S, T>(pop_swap, geo, x, y, z, fo);
    if (WF) {
      u_out[g] = (double)u[0];
      u_out[g + geo.V] = (double)u[1];
      if (S::D == 3) u_out[g + 2 * geo.V] = (double)u[2];
    }
  }
}

// K_ComputeVelocityAfterForceSCMPMeta: u = sum xi f / n + F/(2n)
// (MultiPhaseKernelsMeta.py:1077-1219; with walls the Walls variant,
// :1221-1365, fluid sites only), evaluated in double.
template <class S, class ES, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_velocity(const T* __restrict__ pop, const double* __restrict__ n_in,
           const double* __restrict__ psi_in, Geom geo, double G,
           double* __restrict__ u_out, const uint8_t* __restrict__ walls, double bx, double by,
           double bz) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (walls && walls[g] != 1) continue;
    int x, y, z;
    site_xyz(geo, g, x, y, z);
    double gx, gy, gz;
    grad_global<S, ES, double>(psi_in, geo, x, y, z, gx, gy, gz);
    double j[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const double f = (double)pop[(size_t)q * geo.V + g];
      j[0] += S::cx[q] * f;
      j[1] += S::cy[q] * f;
      j[2] += S::cz[q] * f;
    }
    const double ln = n_in[g];
    const double mGpsi = -G * psi_in[g];
    const double F[3] = {mGpsi * gx + bx, mGpsi * gy + by, mGpsi * gz + bz};
    for (int d = 0; d < S::D; d++) u_out[g + d * geo.V] = j[d] / ln + 0.5 * F[d] / ln;
  }
}

// K_InitPopulations: f = f_eq(n, u) (idpy/LBM/LBM.py:1700-1743).
template <class S, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_init_pop(T* __restrict__ pop, const double* __restrict__ n,
           const double* __restrict__ u, Geom geo, double rho0) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const double ln = n[g];
    const double ux = u[g], uy = u[g + geo.V], uz = S::D == 3 ? u[g + 2 * geo.V] : 0.0;
    const double uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const double xu = S::cx[q] * ux + S::cy[q] * uy + S::cz[q] * uz;
      double e = 1.0;
      e += xu * kCM2;
      e += 0.5 * xu * xu * kCM4;
      e -= 0.5 * uu * kCM2;
      double f = e * ln * S::w[q];
      if (SHIFT) f -= S::w[q] * rho0;
      pop[(size_t)q * geo.V + g] = (T)f;
    }
  }
}

// f_eq from coefficients on the basis coordinates: f_q = sum_k tab[k Q + q]
// c_k, c_k the coordinates of basis_coords (no force). The tables of the
// printed-constant initialisation (K_InitPopulationsMeta with its 15-digit
// literals, idpy/LBM/LBMKernelsMeta.py:282-390) are passed by value.
template <int NBQ>
struct EqTab {
  double c[NBQ];
};

template <class S, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_init_pop_tab(T* __restrict__ pop, const double* __restrict__ n, const double* __restrict__ u,
               Geom geo, double rho0, EqTab<NBasis<S::D>::value * S::Q> tab) {
  constexpr int NB = NBasis<S::D>::value;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const double ln = n[g];
    const double ux = u[g], uy = u[g + geo.V], uz = S::D == 3 ? u[g + 2 * geo.V] : 0.0;
    const double uu = ux * ux + uy * uy + uz * uz;
    double c[NB], d[NB];
    basis_coords<S, double, false>(ln, ln, ux, uy, uz, 0.0, 0.0, 0.0, uu, 0.0, c, d);
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      double f = tab.c[q] * c[0];
#pragma unroll
      for (int k = 1; k < NB; k++) f += tab.c[k * S::Q + q] * c[k];
      if (SHIFT) f -= S::w[q] * rho0;
      pop[(size_t)q * geo.V + g] = (T)f;
    }
  }
}

// Device population layout <-> float64 SoA (adds / removes the shift).
template <class S, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_pop_to_f64(const T* __restrict__ pop, double* __restrict__ dst, Geom geo, double rho0) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      double v = (double)pop[(size_t)q * geo.V + g];
      if (SHIFT) v += S::w[q] * rho0;
      dst[(size_t)q * geo.V + g] = v;
    }
  }
}

template <class S, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_pop_from_f64(T* __restrict__ pop, const double* __restrict__ src, Geom geo, double rho0) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      double v = src[(size_t)q * geo.V + g];
      if (SHIFT) v -= S::w[q] * rho0;
      pop[(size_t)q * geo.V + g] = (T)v;
    }
  }
}

// --------------------------------------------------------------------------
// FP64 reductions: wavefront (64-lane) shuffles, then LDS across waves.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Block of 256 threads = 4 waves. Returns the block total in thread 0.
template <int OP>
__device__ inline double block_reduce(double v) {
  __shared__ double part[4];
  v = OP == 0 ? wave_sum(v) : wave_max(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) part[wv] = v;
  __syncthreads();
  double r = part[0];
  for (int k = 1; k < 4; k++) r = OP == 0 ? r + part[k] : fmax(r, part[k]);
  __syncthreads();
  return r;
}

// Per-block mass and momentum partials of a population buffer:
// out[b*4 + 0] = sum n, out[b*4 + 1 + d] = sum j_d.
template <class S, class T, bool SHIFT>
__global__ void __launch_bounds__(256)
k_mass_mom_partial(const T* __restrict__ pop, Geom geo, double rho0,
                   double* __restrict__ out) {
  double m = 0, jx = 0, jy = 0, jz = 0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < geo.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const double f = (double)pop[(size_t)q * geo.V + g];
      s += f;
      jx += S::cx[q] * f;
      jy += S::cy[q] * f;
      jz += S::cz[q] * f;
    }
    m += SHIFT ? s + rho0 : s;
  }
  m = block_reduce<0>(m);
  jx = block_reduce<0>(jx);
  jy = block_reduce<0>(jy);
  jz = block_reduce<0>(jz);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 4 + 0] = m;
    out[blockIdx.x * 4 + 1] = jx;
    out[blockIdx.x * 4 + 2] = jy;
    out[blockIdx.x * 4 + 3] = jz;
  }
}

}  // namespace idmx
