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

// ------------------------------------------------------------------ Rys roots
#include "rys_table.inc"
#include "rys_table_f32.inc"

template <int N, typename R>
__device__ __forceinline__ void rys_roots(R T, const double* __restrict__ tab, const double* __restrict__ herm,
                                          R* x, R* w) {
  if (T >= (R)RYS_TASYM[N]) {
    R it = (R)1 / T, rs = (R)1 / sqrt(T);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x[i] = (R)herm[RYS_HOFF[N] + i] * it;
      w[i] = (R)herm[RYS_HOFF[N] + N + i] * rs;
    }
    return;
  }
  int k = (int)T;
  R s = (R)2 * (T - (R)k) - (R)1;
  const double* c = tab + RYS_OFF[N] + k * (2 * N * (RYS_DEG + 1));
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    R v = (R)c[f * (RYS_DEG + 1) + RYS_DEG];
#pragma unroll
    for (int j = RYS_DEG - 1; j >= 0; --j) v = v * s + (R)c[f * (RYS_DEG + 1) + j];
    if (f < N) x[f] = v; else w[f - N] = v;
  }
}

// FP32 roots: degree-6 fits, hardware reciprocal / reciprocal square root
template <int N>
__device__ __forceinline__ void rys_roots32(float T, const float* __restrict__ tab, const double* __restrict__ herm,
                                            float* x, float* w) {
  if (T >= (float)RYS_TASYM[N]) {
    const float it = __builtin_amdgcn_rcpf(T), rs = __builtin_amdgcn_rsqf(T);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x[i] = (float)herm[RYS_HOFF[N] + i] * it;
      w[i] = (float)herm[RYS_HOFF[N] + N + i] * rs;
    }
    return;
  }
  const int k = (int)T;
  const float s = 2.0f * (T - (float)k) - 1.0f;
  const float* c = tab + RYS32_OFF[N] + k * (2 * N * (RYS32_DEG + 1));
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    float v = c[f * (RYS32_DEG + 1) + RYS32_DEG];
#pragma unroll
    for (int j = RYS32_DEG - 1; j >= 0; --j) v = v * s + c[f * (RYS32_DEG + 1) + j];
    if (f < N) x[f] = v; else w[f - N] = v;
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
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr long long flops_prim() {
  constexpr int LAB = LA + LB, LCD = LC + LD, nr = (LAB + LCD) / 2 + 1;
  constexpr int nc = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  return (long long)nr * (56 + 12 * (LAB + 1) * (LCD + 1) + 6 * (LA + 1) * (LB + 1) * (LCD + 1) +
                          6 * (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1) + 3 * nc);
}

template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ bool rys_accumulate(const double* __restrict__ Pa, const double* __restrict__ Cc,
                                               const PrimPair& pp, const PrimPair& qq, R qcut2,
                                               const double* __restrict__ tab, const float* __restrict__ tab32,
                                               const double* __restrict__ herm,
                                               R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF]) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int LAB = Dm::LAB, LCD = Dm::LCD, NR = Dm::NROOT;
  constexpr bool F32 = sizeof(R) == 4;
  const R zeta = (R)pp.g, eta = (R)qq.g;
  const R fqz = (R)pp.Kg * (R)qq.Kg;
  const R ze = zeta + eta;
  if (fqz * fqz < qcut2 * ze) return false;
  R ize, sq;
  if constexpr (F32) {
    ize = __builtin_amdgcn_rcpf(ze);
    sq = __builtin_amdgcn_rsqf(ze);
  } else {
    ize = (R)1 / ze;
    sq = sqrt(ize);
  }
  const R rho = zeta * eta * ize;
  R PQ[3], PA[3], QC[3];
  R pq2 = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    PQ[d] = (R)pp.P[d] - (R)qq.P[d];
    PA[d] = (R)pp.P[d] - (R)Pa[d];
    QC[d] = (R)qq.P[d] - (R)Cc[d];
    pq2 += PQ[d] * PQ[d];
  }
  const R T = rho * pq2;
  const R pref = fqz * sq;
  R rx[NR], rw[NR];
  if constexpr (F32) rys_roots32<NR>(T, tab32, herm, rx, rw);
  else rys_roots<NR, R>(T, tab, herm, rx, rw);
  // rho/zeta = eta/(zeta+eta), rho/eta = zeta/(zeta+eta): no further divisions
  const R rz = eta * ize, re = zeta * ize;
  const R h2z = (R)pp.h, h2e = (R)qq.h, h2ze = (R)0.5 * ize;
  // large classes: keep the root loop rolled so the unrolled body stays compilable
  constexpr int UR = (Dm::NE * Dm::NF > 256) ? 1 : NR;
#pragma unroll UR
  for (int r = 0; r < NR; ++r) {
    const R t2 = rx[r];
    const R B00 = h2ze * t2;
    const R B10 = h2z - h2z * rz * t2;
    const R B01 = h2e - h2e * re * t2;
    R G[3][LAB + 1][LCD + 1];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R C00 = PA[d] - rz * t2 * PQ[d];
      const R D00 = QC[d] + re * t2 * PQ[d];
      G[d][0][0] = (d == 0) ? pref * rw[r] : (R)1;
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
#pragma unroll
    for (int e = 0; e < Dm::NE; ++e) {
      const int Le = range_level<LA>(e), ce = range_local<LA>(e);
      const int ex = c_lx(Le, ce), ey = c_ly(Le, ce), ez = c_lz(Le, ce);
#pragma unroll
      for (int f = 0; f < Dm::NF; ++f) {
        const int Lf = range_level<LC>(f), cf = range_local<LC>(f);
        const int fx = c_lx(Lf, cf), fy = c_ly(Lf, cf), fz = c_lz(Lf, cf);
        acc[e][f] += G[0][ex][fx] * G[1][ey][fy] * G[2][ez][fz];
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------- HRR
// Y[a][b][f] = sum_j prod_d C(b_d,j_d) AB_d^(b_d-j_d) acc[(a+j)][f]
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
#pragma unroll
  for (int a = 0; a < ncart(LA); ++a)
#pragma unroll
    for (int b = 0; b < ncart(LB); ++b) {
      const int ax = c_lx(LA, a), ay = c_ly(LA, a), az = c_lz(LA, a);
      const int bx = c_lx(LB, b), by = c_ly(LB, b), bz = c_lz(LB, b);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        R s = 0;
#pragma unroll
        for (int jx = 0; jx <= bx; ++jx)
#pragma unroll
          for (int jy = 0; jy <= by; ++jy)
#pragma unroll
            for (int jz = 0; jz <= bz; ++jz) {
              const int L = LA + jx + jy + jz;
              const int e = range_index<LA>(L, c_idx(ax + jx, ay + jy, az + jz));
              const R cf = (R)(binom(bx, jx) * binom(by, jy) * binom(bz, jz)) * pw[0][bx - jx] *
                           pw[1][by - jy] * pw[2][bz - jz];
              s += cf * acc[e][f];
            }
        Y[a * ncart(LB) + b][f] = s;
      }
    }
}

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
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
    for (int c = 0; c < ncart(LC); ++c)
#pragma unroll
      for (int dd = 0; dd < ncart(LD); ++dd) {
        const int cx = c_lx(LC, c), cy = c_ly(LC, c), cz = c_lz(LC, c);
        const int dx = c_lx(LD, dd), dy = c_ly(LD, dd), dz = c_lz(LD, dd);
        R s = 0;
#pragma unroll
        for (int jx = 0; jx <= dx; ++jx)
#pragma unroll
          for (int jy = 0; jy <= dy; ++jy)
#pragma unroll
            for (int jz = 0; jz <= dz; ++jz) {
              const int L = LC + jx + jy + jz;
              const int f = range_index<LC>(L, c_idx(cx + jx, cy + jy, cz + jz));
              const R cf = (R)(binom(dx, jx) * binom(dy, jy) * binom(dz, jz)) * pw[0][dx - jx] *
                           pw[1][dy - jy] * pw[2][dz - jz];
              s += cf * Y[ab][f];
            }
        V[(ab * ncart(LC) + c) * ncart(LD) + dd] = s;
      }
}

}  // namespace oqph
