// grad_device.h -- the adjoint of the J/K build: derivative ERIs contracted
// with the RHF two-particle density, giving dE2/dR (source/integrals/grd2.F90:
// 145-391 loops and screens, source/integrals/grd2_rys.F90:303-937 integral
// derivatives, source/modules/hf_gradient.F90:428-495 density).
//
// One thread evaluates one canonical shell quartet; a wavefront takes a chunk
// of 64 kets of one bra pair (WaveWork), so the bra pair, its primitives and
// its density block are wave-uniform scalar loads.  Gamma is formed in raw
// (x^l-normalised, canonical-order) Cartesians from the Cartesian-effective
// density Dc = X^T D X,
//   Gamma_abcd = 4 cs Dc_ab Dc_cd - hs (Dc_ac Dc_bd + Dc_ad Dc_bc),
// and contracted per primitive quartet and Rys root with the derivative 2D
// integrals d/dA_x G = 2 alpha G(a+1) - a G(a-1) (likewise B, C); the D
// centre follows from translational invariance.  Per-atom sums are staged in
// LDS and flushed with one FP64 atomic per block and coordinate.
//
// R = float: the mixed-precision form for the quartets the host assigns to
// FP32 (Q_ij Q_kl below the gradient's FP32 cutoff): Gamma, the quadrature
// (FP32 root fits, hardware rsqrt) and the per-quartet derivative sums in
// FP32, the screens and the per-atom sums in FP64.  Its smaller register file
// removes the spills of the d classes' FP64 form.
#pragma once
#include "eri_device.h"
#include "grad_params.h"

namespace oqph {

// wave-aggregated counters (every lane of the wave calls them)
__device__ __forceinline__ void sum_wave_g(unsigned long long* ctr, unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (v != 0ull && __lane_id() == 0) atomicAdd(ctr, v);
}
__device__ __forceinline__ void count_wave_g(unsigned long long* ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m != 0ull && __lane_id() == 0) atomicAdd(ctr, (unsigned long long)__popcll(m));
}



// 2D integral with bra/ket horizontal transfer, all indices compile-time:
//   I(a,b,c,d) = sum_j sum_m C(b,j) AB^(b-j) C(d,m) CD^(d-m) G(a+j, c+m)
template <int A, int B, int C, int D, int NI, int NK, typename R>
__device__ __forceinline__ R i2d(const R (&G)[NI][NK], const R* pab, const R* pcd) {
  R s = 0;
  sfor<B + 1>([&](auto Ji) {
    constexpr int j = OQPH_CI(Ji);
    sfor<D + 1>([&](auto Mi) {
      constexpr int m = OQPH_CI(Mi);
      constexpr int cb = binom(B, j) * binom(D, m);
      s += (R)cb * pab[B - j] * pcd[D - m] * G[A + j][C + m];
    });
  });
  return s;
}

template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr long long flops_grad_prim() {
  return flops_grad_prim_rt(LA, LB, LC, LD);
}

// MRSF: Gamma from the MRSF matrices (prm.Mc) instead of the density
template <int LA, int LB, int LC, int LD, typename R, bool MRSF = false>
__global__ __launch_bounds__(256) void k_grad(GradParams prm, WaveWork ww) {
  extern __shared__ double s_de[];
  constexpr bool F32 = sizeof(R) == 4;
  constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
  // the root table block of this class in LDS when it fits (per-lane gathers
  // at random intervals are cheaper from LDS than through vector memory)
  constexpr bool T64 = !F32 && rys_len<NR>() * 8 <= 36 * 1024;
  constexpr bool T32 = F32;
  __shared__ double s_t64[T64 ? rys_len<NR>() : 1];
  __shared__ float s_t32[T32 ? rys32_len<NR>() : 1];
  for (int i = threadIdx.x; i < 3 * prm.natm; i += blockDim.x) s_de[i] = 0.0;
  if constexpr (T64)
    for (int i = threadIdx.x; i < rys_len<NR>(); i += blockDim.x) s_t64[i] = prm.rys[RYS_OFF[NR] + i];
  if constexpr (T32)
    for (int i = threadIdx.x; i < rys32_len<NR>(); i += blockDim.x) s_t32[i] = prm.rys32[RYS32_OFF[NR] + i];
  __syncthreads();
  const double* tabN = T64 ? (const double*)s_t64 : prm.rys + RYS_OFF[NR];
  const float* tab32N = (const float*)s_t32;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NG = NA * NB * NC * ND;
  constexpr int LAB = LA + LB + 1, LCD = LC + LD + 1;
  // one wavefront = one chunk of 64 kets of one segment; the bra is uniform.
  // Waves past the end still reach the block barrier below.
  const long long wave = __builtin_amdgcn_readfirstlane(
      (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6));
  const bool wave_ok = wave < ww.nwaves;
  bool active = false;
  bool skip = true;
  ShellPair P{}, Q{};
  int ip = 0, iq = 0;
  R Gam[NG];
  double gm = 0.0;
  if (wave_ok) {
    int lo = 0, hi = ww.nseg;
    while (hi - lo > 1) {  // scalar binary search over the chunk prefix
      const int mid = (lo + hi) >> 1;
      if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
    }
    const int seg = __builtin_amdgcn_readfirstlane(lo);
    const int k = (int)(wave - ww.choff[seg]) * 64 + (int)__lane_id();
    active = k < ww.seg_n[seg];
    ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
    P = prm.pairs[ip];
    if (active) iq = ww.ket[ww.seg_k0[seg] + k];
  }
  if (active) {
    Q = prm.pairs[iq];
    const int oa = prm.coff[P.sa], ob = prm.coff[P.sb], oc = prm.coff[Q.sa], od = prm.coff[Q.sb];
    const int n = prm.ncf;
    const double* Dc = prm.Dc;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            double g;
            if constexpr (MRSF)
              g = mrsf_gamma(prm, oa + a, ob + b, oc + c, od + d);
            else
              g = 4.0 * prm.cs * Dc[(oa + a) * n + ob + b] * Dc[(oc + c) * n + od + d] -
                  prm.hs * (Dc[(oa + a) * n + oc + c] * Dc[(ob + b) * n + od + d] +
                            Dc[(oa + a) * n + od + d] * Dc[(ob + b) * n + oc + c]);
            Gam[((a * NB + b) * NC + c) * ND + d] = (R)g;
            gm = fmax(gm, fabs(g));
          }
    // fine screen on the weighted contribution (grd2.F90:348)
    skip = gm * P.Q * Q.Q < 0.5 * prm.gcut;
  }
  unsigned long long* ctr = prm.counters + 8 * (int)(wave & (OQPH_CTR_SLOTS - 1));
  count_wave_g(ctr + 1, active && skip);
  count_wave_g(ctr + 0, !skip);
  R fd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int nprim = 0;
  if (!skip) {
    const double dab2 = gm * gm;
    for (int iqq = 0; iqq < Q.nprim; ++iqq) {
      const PrimPair qq = prm.prims[Q.prim0 + iqq];
      for (int ipp = 0; ipp < P.nprim; ++ipp) {
        const PrimPair pp = load_uniform(prm.prims, P.prim0 + ipp);
        const double zeta = pp.g, eta = qq.g, ze = zeta + eta;
        const double fqz = pp.Kg * qq.Kg;
        // primitive screens (grd2_rys.F90:375-376), prefactor cutoff (1e-20)^2
        if (fqz * fqz < 1.0e-40 * ze) continue;
        if (fqz * fqz * dab2 < prm.dabcut2 * ze) continue;
        ++nprim;
        // differences of absolute positions in FP64, then the working precision
        R PQ[3], PA[3], QC[3], pq2 = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          PQ[d] = (R)(pp.P[d] - qq.P[d]);
          PA[d] = (R)(pp.P[d] - P.A[d]);
          QC[d] = (R)(qq.P[d] - Q.A[d]);
          pq2 += PQ[d] * PQ[d];
        }
        const R zr = (R)zeta, er = (R)eta;
        const R sq = F32 ? (R)__builtin_amdgcn_rsqf((float)(zr + er)) : (R)rsq64(ze), ize = sq * sq,
                rho = zr * er * ize;
        R rx[NR], rw[NR];
        if constexpr (F32) rys_roots32<NR>((float)(rho * pq2), tab32N, rx, rw);
        else rys_roots<NR>(rho * pq2, tabN, rx, rw);
        const R pref = (R)fqz * sq;
        const R rz = er * ize, re = zr * ize;
        const R h2z = (R)pp.h, h2e = (R)qq.h, h2ze = (R)0.5 * ize;
        const R a2 = (R)(2.0 * pp.a), b2 = (R)(2.0 * pp.b), c2 = (R)(2.0 * qq.a);
        constexpr int UR = (NG > 36) ? 1 : NR;
#pragma unroll UR
        for (int r = 0; r < NR; ++r) {
          const R t2 = rx[r];
          const R B00 = h2ze * t2, B10 = h2z - h2z * rz * t2, B01 = h2e - h2e * re * t2;
          R G[3][LAB + 1][LCD + 1];
          R pab[3][LB + 2], pcd[3][LD + 1];
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
            for (int k = 1; k <= LCD; ++k)
#pragma unroll
              for (int i = 0; i <= LAB; ++i) {
                R v = D00 * G[d][i][k - 1];
                if (k > 1) v += (R)(k - 1) * B01 * G[d][i][k - 2];
                if (i > 0) v += (R)i * B00 * G[d][i - 1][k - 1];
                G[d][i][k] = v;
              }
            const R ab = (R)(P.A[d] - P.B[d]), cd = (R)(Q.A[d] - Q.B[d]);
            pab[d][0] = (R)1;
#pragma unroll
            for (int k = 1; k <= LB + 1; ++k) pab[d][k] = pab[d][k - 1] * ab;
            pcd[d][0] = (R)1;
#pragma unroll
            for (int k = 1; k <= LD; ++k) pcd[d][k] = pcd[d][k - 1] * cd;
          }
          sfor<NA>([&](auto Ai) {
            constexpr int a = OQPH_CI(Ai);
            sfor<NB>([&](auto Bi) {
              constexpr int b = OQPH_CI(Bi);
              sfor<NC>([&](auto Ci) {
                constexpr int c = OQPH_CI(Ci);
                sfor<ND>([&](auto Di) {
                  constexpr int dd = OQPH_CI(Di);
                  const R g = Gam[((a * NB + b) * NC + c) * ND + dd];
                  R I0[3], dA[3], dB[3], dC[3];
                  sfor<3>([&](auto Xi) {
                    constexpr int x = OQPH_CI(Xi);
                    constexpr int ia = c_l(LA, a, x), ib = c_l(LB, b, x), ic = c_l(LC, c, x), id = c_l(LD, dd, x);
                    const R* pa = pab[x];
                    const R* pc = pcd[x];
                    I0[x] = i2d<ia, ib, ic, id>(G[x], pa, pc);
                    R v = a2 * i2d<ia + 1, ib, ic, id>(G[x], pa, pc);
                    if constexpr (ia > 0) v -= (R)ia * i2d<ia - 1, ib, ic, id>(G[x], pa, pc);
                    dA[x] = v;
                    v = b2 * i2d<ia, ib + 1, ic, id>(G[x], pa, pc);
                    if constexpr (ib > 0) v -= (R)ib * i2d<ia, ib - 1, ic, id>(G[x], pa, pc);
                    dB[x] = v;
                    v = c2 * i2d<ia, ib, ic + 1, id>(G[x], pa, pc);
                    if constexpr (ic > 0) v -= (R)ic * i2d<ia, ib, ic - 1, id>(G[x], pa, pc);
                    dC[x] = v;
                  });
                  fd[0][0] += g * dA[0] * I0[1] * I0[2];
                  fd[0][1] += g * I0[0] * dA[1] * I0[2];
                  fd[0][2] += g * I0[0] * I0[1] * dA[2];
                  fd[1][0] += g * dB[0] * I0[1] * I0[2];
                  fd[1][1] += g * I0[0] * dB[1] * I0[2];
                  fd[1][2] += g * I0[0] * I0[1] * dB[2];
                  fd[2][0] += g * dC[0] * I0[1] * I0[2];
                  fd[2][1] += g * I0[0] * dC[1] * I0[2];
                  fd[2][2] += g * I0[0] * I0[1] * dC[2];
                });
              });
            });
          });
        }
      }
    }
  }
  sum_wave_g(ctr + 3, (unsigned long long)nprim * flops_grad_prim<LA, LB, LC, LD>());
  sum_wave_g(ctr + 4, (unsigned long long)nprim);
  if (!skip && nprim > 0) {
    double deg = 1.0;
    if (P.sa == P.sb) deg *= 0.5;
    if (Q.sa == Q.sb) deg *= 0.5;
    if (ip == iq) deg *= 0.5;
    const int at[4] = {prm.shell_atom[P.sa], prm.shell_atom[P.sb], prm.shell_atom[Q.sa], prm.shell_atom[Q.sb]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double f0 = (double)fd[0][d], f1 = (double)fd[1][d], f2 = (double)fd[2][d];
      const double fD = -(f0 + f1 + f2);
      atomicAdd(&s_de[3 * at[0] + d], deg * f0);
      atomicAdd(&s_de[3 * at[1] + d], deg * f1);
      atomicAdd(&s_de[3 * at[2] + d], deg * f2);
      atomicAdd(&s_de[3 * at[3] + d], deg * fD);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * prm.natm; i += blockDim.x)
    if (s_de[i] != 0.0) unsafeAtomicAdd(prm.de + (size_t)(blockIdx.x & (OQPH_DE_SLOTS - 1)) * 3 * prm.natm + i, s_de[i]);
}

}  // namespace oqph
