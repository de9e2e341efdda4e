// eri_device.h -- device-side building blocks of the J/K and gradient kernels.
//
// Electron-repulsion integrals are evaluated by Rys quadrature (Dupuis, Rys,
// King, J. Chem. Phys. 65, 111 (1976)): per primitive quartet and root, the
// 2D integrals G_d(i,k), i <= La+Lb, k <= Lc+Ld, come from the vertical
// recursion on centres A and C; the Cartesian (e0|f0) products are contracted
// over primitives and roots in registers; the horizontal transfer to (ab|cd)
// runs once per contracted quartet in its closed binomial form
//   (a b| = sum_j prod_d C(b_d, j_d) AB_d^(b_d - j_d) (a + j, 0| .
// Everything that depends on the angular momenta is a compile-time constant,
// so the loops unroll into straight-line register code for each class.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>
#include "oqph_common.h"

namespace oqph {

// ---------------------------------------------------------------- Cartesian
__host__ __device__ constexpr int ncart(int L) { return L < 0 ? 0 : (L + 1) * (L + 2) / 2; }
__host__ __device__ constexpr int ncum(int L) { return L < 0 ? 0 : (L + 1) * (L + 2) * (L + 3) / 6; }
// canonical order inside level L: lx descending, then ly descending
__host__ __device__ constexpr int c_k(int c) {  // k = L - lx
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  return k;
}
__host__ __device__ constexpr int c_lx(int L, int c) { return L - c_k(c); }
__host__ __device__ constexpr int c_lz(int L, int c) { return c - c_k(c) * (c_k(c) + 1) / 2; }
__host__ __device__ constexpr int c_ly(int L, int c) { return c_k(c) - c_lz(L, c); }
__host__ __device__ constexpr int c_idx(int lx, int ly, int lz) {  // local index in level lx+ly+lz
  return (ly + lz) * (ly + lz + 1) / 2 + lz;
}
// component of level L along axis d
__host__ __device__ constexpr int c_l(int L, int c, int d) {
  return d == 0 ? c_lx(L, c) : (d == 1 ? c_ly(L, c) : c_lz(L, c));
}
__host__ __device__ constexpr int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - i + 1) / i;
  return r;
}

// ------------------------------------------------------ compile-time loops
// sfor<N>(f) calls f(std::integral_constant<int, i>{}) for i = 0..N-1, so the
// body sees every index as a constant expression: Cartesian exponents, HRR
// binomials and register-array offsets are all resolved at compile time.
template <typename F, int... I>
__device__ __forceinline__ void sfor_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (N > 0) sfor_impl(f, std::make_integer_sequence<int, N>{});
}
#define OQPH_CI(x) decltype(x)::value

// ------------------------------------------------------------------ FP64 rcp
// Hardware reciprocal / reciprocal square root seeds refined by two Newton
// steps (quadratic convergence from the seed's ~2^-22 to full precision);
// arguments are positive normal numbers (exponent sums, Boys arguments).
__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double rsq64(double x) {
  const double h = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
}

// Load a primitive pair whose index is uniform across the wavefront through
// the constant address space, so it is served by the scalar cache.
__device__ __forceinline__ PrimJK load_uniform(const PrimJK* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(4))) double* q =
      (const __attribute__((address_space(4))) double*)(p + i);
  PrimJK r;
  r.g = q[0]; r.Kg = q[1]; r.X[0] = q[2]; r.X[1] = q[3]; r.X[2] = q[4]; r.h = q[5];
  return r;
#else
  return p[i];
#endif
}
__device__ __forceinline__ PrimF32 load_uniform(const PrimF32* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(4))) float* q =
      (const __attribute__((address_space(4))) float*)(p + i);
  PrimF32 r;
  r.g = q[0]; r.Kg = q[1]; r.X[0] = q[2]; r.X[1] = q[3]; r.X[2] = q[4]; r.h = q[5];
  return r;
#else
  return p[i];
#endif
}
__device__ __forceinline__ PrimPair load_uniform(const PrimPair* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(4))) double* q =
      (const __attribute__((address_space(4))) double*)(p + i);
  PrimPair r;
  r.g = q[0]; r.Kg = q[1]; r.P[0] = q[2]; r.P[1] = q[3]; r.P[2] = q[4]; r.a = q[5]; r.b = q[6]; r.h = q[7];
  return r;
#else
  return p[i];
#endif
}

// ------------------------------------------------------------------ Rys roots
#include "rys_table.inc"
#include "rys_table_f32.inc"

// length of the N-root block of each table (entries); the blocks start at
// RYS_OFF[N] / RYS32_OFF[N]
template <int N>
__host__ __device__ constexpr int rys_len() { return RYS_TASYM[N] * 2 * N * (RYS_DEG + 1); }
template <int N>
__host__ __device__ constexpr int rys32_len() { return RYS_TASYM[N] * 2 * N * (RYS32_DEG + 1); }

// Nodes t^2 and weights of the N-root Rys quadrature at argument T.  tabN is
// the N-root block of the FP64 table (global memory or a block-shared copy):
// per unit interval k = floor(T), 2N degree-12 polynomials in s = 2(T-k)-1.
// Beyond RYS_TASYM[N] the Hermite limit t^2 = r_i / T, w = h_i / sqrt(T).
template <int N>
__device__ __forceinline__ void rys_roots(double T, const double* __restrict__ tabN, double* x, double* w) {
  if (T >= (double)RYS_TASYM[N]) {
    const double rs = rsq64(T), it = rs * rs;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x[i] = RYS_HERM[RYS_HOFF[N] + i] * it;
      w[i] = RYS_HERM[RYS_HOFF[N] + N + i] * rs;
    }
    return;
  }
  const int k = (int)T;
  const double s = 2.0 * (T - (double)k) - 1.0;
  const double* c = tabN + k * (2 * N * (RYS_DEG + 1));
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    double v = c[f * (RYS_DEG + 1) + RYS_DEG];
#pragma unroll
    for (int j = RYS_DEG - 1; j >= 0; --j) v = fma(v, s, c[f * (RYS_DEG + 1) + j]);
    if (f < N) x[f] = v; else w[f - N] = v;
  }
}

// FP32 roots: degree-6 fits on the same intervals, hardware rcp / rsqrt
template <int N>
__device__ __forceinline__ void rys_roots32(float T, const float* __restrict__ tabN, float* x, float* w) {
  if (T >= (float)RYS_TASYM[N]) {
    const float rs = __builtin_amdgcn_rsqf(T), it = rs * rs;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x[i] = (float)RYS_HERM[RYS_HOFF[N] + i] * it;
      w[i] = (float)RYS_HERM[RYS_HOFF[N] + N + i] * rs;
    }
    return;
  }
  const int k = (int)T;
  const float s = 2.0f * (T - (float)k) - 1.0f;
  const float* c = tabN + k * (2 * N * (RYS32_DEG + 1));
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    float v = c[f * (RYS32_DEG + 1) + RYS32_DEG];
#pragma unroll
    for (int j = RYS32_DEG - 1; j >= 0; --j) v = fmaf(v, s, c[f * (RYS32_DEG + 1) + j]);
    if (f < N) x[f] = v; else w[f - N] = v;
  }
}

// One node/weight pair (root r, runtime index) of the N-root quadrature, for
// root loops kept rolled to bound register pressure.
template <int N, typename R>
__device__ __forceinline__ void rys_root_one(R T, const double* __restrict__ tabN, const float* __restrict__ tab32N,
                                             int r, R& x, R& w) {
  if constexpr (sizeof(R) == 4) {
    if (T >= (float)RYS_TASYM[N]) {
      const float rs = __builtin_amdgcn_rsqf(T);
      x = (float)RYS_HERM[RYS_HOFF[N] + r] * (rs * rs);
      w = (float)RYS_HERM[RYS_HOFF[N] + N + r] * rs;
      return;
    }
    const int k = (int)T;
    const float s = 2.0f * (T - (float)k) - 1.0f;
    const float* cx = tab32N + k * (2 * N * (RYS32_DEG + 1)) + r * (RYS32_DEG + 1);
    const float* cw = cx + N * (RYS32_DEG + 1);
    float vx = cx[RYS32_DEG], vw = cw[RYS32_DEG];
#pragma unroll
    for (int j = RYS32_DEG - 1; j >= 0; --j) {
      vx = fmaf(vx, s, cx[j]);
      vw = fmaf(vw, s, cw[j]);
    }
    x = vx;
    w = vw;
  } else {
    if (T >= (double)RYS_TASYM[N]) {
      const double rs = rsq64(T);
      x = RYS_HERM[RYS_HOFF[N] + r] * rs * rs;
      w = RYS_HERM[RYS_HOFF[N] + N + r] * rs;
      return;
    }
    const int k = (int)T;
    const double s = 2.0 * (T - (double)k) - 1.0;
    const double* cx = tabN + k * (2 * N * (RYS_DEG + 1)) + r * (RYS_DEG + 1);
    const double* cw = cx + N * (RYS_DEG + 1);
    double vx = cx[RYS_DEG], vw = cw[RYS_DEG];
#pragma unroll
    for (int j = RYS_DEG - 1; j >= 0; --j) {
      vx = fma(vx, s, cx[j]);
      vw = fma(vw, s, cw[j]);
    }
    x = vx;
    w = vw;
  }
}

// ------------------------------------------------- contracted (e0|f0) via Rys
// acc[e][f]: e over levels LA..LA+LB (local offset ncum(LA-1)), f over LC..LC+LD.
template <int LA, int LB, int LC, int LD>
struct Dims {
  static constexpr int LAB = LA + LB, LCD = LC + LD;
  static constexpr int NE = ncum(LAB) - ncum(LA - 1);
  static constexpr int NF = ncum(LCD) - ncum(LC - 1);
  static constexpr int NROOT = (LAB + LCD) / 2 + 1;
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NBLK = NA * NB * NC * ND;
};

// level and local index of the e-th entry of a level range starting at L0
template <int L0>
__host__ __device__ constexpr int range_level(int e) {
  int L = L0;
  while (e >= ncart(L)) { e -= ncart(L); ++L; }
  return L;
}
template <int L0>
__host__ __device__ constexpr int range_local(int e) {
  int L = L0;
  while (e >= ncart(L)) { e -= ncart(L); ++L; }
  return e;
}
template <int L0>
__host__ __device__ constexpr int range_index(int L, int c) { return ncum(L - 1) - ncum(L0 - 1) + c; }

// Algorithmic FLOP model of one surviving primitive quartet (Rys): roots,
// recursion coefficients, 2D VRR, (e0|f0) assembly -- shared with the
// benchmark's roofline accounting.
__host__ __device__ constexpr long long flops_prim_rt(int LA, int LB, int LC, int LD) {
  const int LAB = LA + LB, LCD = LC + LD, nr = (LAB + LCD) / 2 + 1;
  const int nc = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  return (long long)nr * (56 + 12 * (LAB + 1) * (LCD + 1) + 6 * (LA + 1) * (LB + 1) * (LCD + 1) +
                          6 * (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1) + 3 * nc);
}
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr long long flops_prim() {
  return flops_prim_rt(LA, LB, LC, LD);
}
// The same model for the derivative integrals of the gradient: the
// differentiated centres raised by one, one more root level, and the
// assembly term four times (value and three derivative components).
__host__ __device__ constexpr long long flops_grad_prim_rt(int LA, int LB, int LC, int LD) {
  const int LAB = LA + LB + 1, LCD = LC + LD + 1, nr = (LAB + LCD - 1) / 2 + 1;
  const int nc = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  return (long long)nr * (56 + 12 * (LAB + 1) * (LCD + 1) + 6 * (LA + 2) * (LB + 2) * (LCD + 1) +
                          6 * (LA + 2) * (LB + 2) * (LC + 2) * (LD + 1) + 3 * 4 * nc);
}

// The fields of a primitive pair that the quadrature reads, in one precision:
// X is the product centre relative to the pair's first centre.
template <typename R>
struct PrimV {
  R g, Kg, X[3], h;
};
template <typename R, typename PP>
__device__ __forceinline__ PrimV<R> prim_v(const PP& p) {
  return {(R)p.g, (R)p.Kg, {(R)p.X[0], (R)p.X[1], (R)p.X[2]}, (R)p.h};
}

template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void rys_add(const PrimV<R>& pp, const PrimV<R>& qq, const R (&AQ)[3], R ze, R sq, R ize,
                                        R pref, const double* __restrict__ tabN, const float* __restrict__ tab32N,
                                        R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF]);

// The FP32 quadrature of the small classes folds the ket's product centre
// into the centre difference once per ket primitive pair: AQ = (A - C) - X_q,
// and PQ = AQ + X_p in the loop over bra primitives ((ps|ss) mixed 92 -> 89
// ms on (H2O)64/cc-pVDZ, with the ket prefetch).  Elsewhere AQ = A - C and
// PQ = AQ + X_p - X_q per quartet: FP64 needs all of the ket's 48-byte record
// before its bra loop then ((ps|ss) FP64 122 -> 139 ms), and the larger
// classes have no registers to spare.
template <int LA, int LB, int LC, int LD, typename R>
constexpr bool hoist_ket() {
  return sizeof(R) == 4 && Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF <= 9;
}
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void ket_offset(const R (&AC)[3], const PrimV<R>& qq, R (&AQ)[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) AQ[d] = hoist_ket<LA, LB, LC, LD, R>() ? AC[d] - qq.X[d] : AC[d];
}

// Adds one primitive quartet's (e0|f0) to acc; returns false when the
// primitive prefactor screen |K_ab K_cd| / sqrt(zeta+eta) < qcut drops it.
// pp, qq: bra and ket primitive pairs (X = P - A and X = Q - C); AQ from
// ket_offset, so PA = X_p, QC = X_q and PQ = AQ + X_p (- X_q).
// imu2 = 1/mu^2 selects the erf(mu r)/r operator: zeta + eta becomes
// zeta + eta + zeta eta / mu^2 in every quantity of the quadrature
// (OpenQP's attenuated Rys setup, source/integrals/int_rys.F90:180-182, 227);
// imu2 = 0 is the plain Coulomb operator.
// tabN / tab32N: the NROOT-root blocks of the FP64 / FP32 root tables.
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ bool rys_accumulate(const PrimV<R>& pp, const PrimV<R>& qq, const R (&AQ)[3], R qcut2,
                                               R imu2, const double* __restrict__ tabN,
                                               const float* __restrict__ tab32N,
                                               R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF]) {
  constexpr bool F32 = sizeof(R) == 4;
  const R zeta = pp.g, eta = qq.g;
  const R fqz = pp.Kg * qq.Kg;
  const R ze = zeta + eta + zeta * eta * imu2;
  if (fqz * fqz < qcut2 * ze) return false;
  // 1/sqrt(ze) and its square: one transcendental (FP32: hardware rsq)
  const R sq = F32 ? (R)__builtin_amdgcn_rsqf((float)ze) : (R)rsq64((double)ze);
  const R ize = sq * sq;
  rys_add<LA, LB, LC, LD, R>(pp, qq, AQ, ze, sq, ize, fqz * sq, tabN, tab32N, acc);
  return true;
}

// The quadrature of one primitive quartet without the screen: adds the
// (e0|f0) products, scaled by pref, to acc.  ze = zeta + eta (+ zeta eta /
// mu^2 for the attenuated operator), sq = 1/sqrt(ze), ize = 1/ze.
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void rys_add(const PrimV<R>& pp, const PrimV<R>& qq, const R (&AQ)[3], R ze, R sq, R ize,
                                        R pref, const double* __restrict__ tabN, const float* __restrict__ tab32N,
                                        R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF]) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int LAB = Dm::LAB, LCD = Dm::LCD, NR = Dm::NROOT;
  constexpr bool F32 = sizeof(R) == 4;
  const R zeta = pp.g, eta = qq.g;
  const R rho = zeta * eta * ize;
  R PQ[3], PA[3], QC[3];
  R pq2 = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    PA[d] = pp.X[d];
    QC[d] = qq.X[d];
    PQ[d] = hoist_ket<LA, LB, LC, LD, R>() ? AQ[d] + pp.X[d] : AQ[d] + pp.X[d] - qq.X[d];
    pq2 += PQ[d] * PQ[d];
  }
  const R T = rho * pq2;
  // rho/zeta = eta/(zeta+eta), rho/eta = zeta/(zeta+eta): no further divisions
  const R rz = eta * ize, re = zeta * ize;
  const R h2z = pp.h, h2e = qq.h, h2ze = (R)0.5 * ize;
  // one root: 2D integrals G_d(i, k) by the vertical recursion, then the
  // (e0|f0) products
  auto root = [&](const R t2, const R wr) {
    const R B00 = h2ze * t2;
    const R B10 = h2z - h2z * rz * t2;
    const R B01 = h2e - h2e * re * t2;
    R G[3][LAB + 1][LCD + 1];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R C00 = PA[d] - rz * t2 * PQ[d];
      const R D00 = QC[d] + re * t2 * PQ[d];
      G[d][0][0] = (d == 0) ? pref * wr : (R)1;
#pragma unroll
      for (int i = 1; i <= LAB; ++i) {
        R v = C00 * G[d][i - 1][0];
        if (i > 1) v += (R)(i - 1) * B10 * G[d][i - 2][0];
        G[d][i][0] = v;
      }
#pragma unroll
      for (int k = 1; k <= LCD; ++k) {
#pragma unroll
        for (int i = 0; i <= LAB; ++i) {
          R v = D00 * G[d][i][k - 1];
          if (k > 1) v += (R)(k - 1) * B01 * G[d][i][k - 2];
          if (i > 0) v += (R)i * B00 * G[d][i - 1][k - 1];
          G[d][i][k] = v;
        }
      }
    }
    sfor<Dm::NE>([&](auto E) {
      constexpr int e = OQPH_CI(E);
      constexpr int Le = range_level<LA>(e), ce = range_local<LA>(e);
      constexpr int ex = c_lx(Le, ce), ey = c_ly(Le, ce), ez = c_lz(Le, ce);
      sfor<Dm::NF>([&](auto F) {
        constexpr int f = OQPH_CI(F);
        constexpr int Lf = range_level<LC>(f), cf = range_local<LC>(f);
        constexpr int fx = c_lx(Lf, cf), fy = c_ly(Lf, cf), fz = c_lz(Lf, cf);
        acc[e][f] += G[0][ex][fx] * G[1][ey][fy] * G[2][ez][fz];
      });
    });
  };
  if constexpr (Dm::NE * Dm::NF <= 64) {
    // small classes: all roots at once, root loop unrolled
    R rx[NR], rw[NR];
    if constexpr (F32) rys_roots32<NR>(T, tab32N, rx, rw);
    else rys_roots<NR>(T, tabN, rx, rw);
#pragma unroll
    for (int r = 0; r < NR; ++r) root(rx[r], rw[r]);
  } else {
    // larger classes: one root per iteration of a rolled loop, so only one
    // root's 2D integrals are live next to the accumulators
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
      R t2, wr;
      rys_root_one<NR, R>(T, tabN, tab32N, r, t2, wr);
      root(t2, wr);
    }
  }
}

// ---------------------------------------------------------------- HRR
// Y[a][b][f] = sum_j prod_d C(b_d,j_d) AB_d^(b_d-j_d) acc[(a+j)][f]
// f runs outermost so each column of acc dies as its Y column is formed.
template <int LA, int LB, int NF, typename R>
__device__ __forceinline__ void hrr_bra(const R (&acc)[ncum(LA + LB) - ncum(LA - 1)][NF], const R* AB,
                                        R (&Y)[ncart(LA) * ncart(LB)][NF]) {
  R pw[3][LB + 1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pw[d][0] = (R)1;
#pragma unroll
    for (int k = 1; k <= LB; ++k) pw[d][k] = pw[d][k - 1] * AB[d];
  }
  sfor<NF>([&](auto Fi) {
    constexpr int f = OQPH_CI(Fi);
    sfor<ncart(LA)>([&](auto Ai) {
      constexpr int a = OQPH_CI(Ai);
      constexpr int ax = c_lx(LA, a), ay = c_ly(LA, a), az = c_lz(LA, a);
      sfor<ncart(LB)>([&](auto Bi) {
        constexpr int b = OQPH_CI(Bi);
        constexpr int bx = c_lx(LB, b), by = c_ly(LB, b), bz = c_lz(LB, b);
        R s = 0;
        sfor<bx + 1>([&](auto JX) {
          constexpr int jx = OQPH_CI(JX);
          sfor<by + 1>([&](auto JY) {
            constexpr int jy = OQPH_CI(JY);
            sfor<bz + 1>([&](auto JZ) {
              constexpr int jz = OQPH_CI(JZ);
              constexpr int e = range_index<LA>(LA + jx + jy + jz, c_idx(ax + jx, ay + jy, az + jz));
              constexpr int cb = binom(bx, jx) * binom(by, jy) * binom(bz, jz);
              s += (R)cb * pw[0][bx - jx] * pw[1][by - jy] * pw[2][bz - jz] * acc[e][f];
            });
          });
        });
        Y[a * ncart(LB) + b][f] = s;
      });
    });
  });
}

// V[ab][c][d] = sum_j prod_d C(d_d,j_d) CD_d^(d_d-j_d) Y[ab][(c+j)]
template <int LC, int LD, int NAB, typename R>
__device__ __forceinline__ void hrr_ket(const R (&Y)[NAB][ncum(LC + LD) - ncum(LC - 1)], const R* CD,
                                        R (&V)[NAB * ncart(LC) * ncart(LD)]) {
  R pw[3][LD + 1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pw[d][0] = (R)1;
#pragma unroll
    for (int k = 1; k <= LD; ++k) pw[d][k] = pw[d][k - 1] * CD[d];
  }
  sfor<NAB>([&](auto ABi) {
    constexpr int ab = OQPH_CI(ABi);
    sfor<ncart(LC)>([&](auto Ci) {
      constexpr int c = OQPH_CI(Ci);
      constexpr int cx = c_lx(LC, c), cy = c_ly(LC, c), cz = c_lz(LC, c);
      sfor<ncart(LD)>([&](auto Di) {
        constexpr int dd = OQPH_CI(Di);
        constexpr int dx = c_lx(LD, dd), dy = c_ly(LD, dd), dz = c_lz(LD, dd);
        R s = 0;
        sfor<dx + 1>([&](auto JX) {
          constexpr int jx = OQPH_CI(JX);
          sfor<dy + 1>([&](auto JY) {
            constexpr int jy = OQPH_CI(JY);
            sfor<dz + 1>([&](auto JZ) {
              constexpr int jz = OQPH_CI(JZ);
              constexpr int f = range_index<LC>(LC + jx + jy + jz, c_idx(cx + jx, cy + jy, cz + jz));
              constexpr int cb = binom(dx, jx) * binom(dy, jy) * binom(dz, jz);
              s += (R)cb * pw[0][dx - jx] * pw[1][dy - jy] * pw[2][dz - jz] * Y[ab][f];
            });
          });
        });
        V[(ab * ncart(LC) + c) * ncart(LD) + dd] = s;
      });
    });
  });
}

}  // namespace oqph
