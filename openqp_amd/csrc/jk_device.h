// jk_kernels.hip -- class-templated J/K Fock-build kernels for gfx950.
//
// One thread evaluates one canonical shell quartet (ab|cd) of a fixed
// angular-momentum class: Rys quadrature over all surviving primitive
// quartets with the contracted (e0|f0) set held in registers, horizontal
// transfer, the transform to OpenQP's AO representation (component order,
// shells_pnrm2 norms, pure projection), OpenQP's element cutoff, and the
// digestion into the packed two-electron Fock matrix
//     G = sc * J[D] - 0.5 * se * K[D]
// with FP64 atomics.  The result equals what fock_jk returns after its
// 0.5 scaling and diagonal doubling (source/scf_addons.F90:1178-1185).
//
// Screening restated from source/integrals/int2.F90:
//   pair test    Q_bra * maxD < c                         (:963-971, :763-770)
//   quartet test Q_ij Q_kl max(4D_ij,4D_kl,D_jl,D_jk,D_il,D_ik) < c   (:975-986)
//   element test |(ab|cd)| < c                           (:1806-1812)
// where "bra" is the pair OpenQP's loop visits first: the one with the larger
// (i, j) shell index, i >= j (:755-780).
#pragma once
#include "eri_device.h"

namespace oqph {

// threads per J/K block: 8 wavefronts share one LDS copy of the root tables
#define JK_BLOCK 512

// Canonical-Cartesian -> AO transforms for L = 2, 3 (Cartesian and pure) live
// in device memory as xform[2][2][10][10]: rows = AO component, cols =
// canonical Cartesian component; built by the host from the solid-harmonic
// formula (session.hip).



__device__ __forceinline__ long long tri_idx(int i, int j) {  // 0-based, any order
  int a = i > j ? i : j, b = i > j ? j : i;
  return (long long)a * (a + 1) / 2 + b;
}

template <int L, bool PURE>
struct AO {
  static constexpr int N = (PURE && L >= 2) ? 2 * L + 1 : ncart(L);
};

// AO transform of one index of a 4-index block, in place: V is viewed as
// [LEFT][NI][RIGHT] on input and [LEFT][NO][RIGHT] on output (NO <= NI).
// Each (l, r) line is read completely before it is written, in increasing l,
// so no write lands on an unread input: within one l the lines differ in r,
// and the outputs of line l end below the first input of line l+1.
template <int L, bool PURE, int LEFT, int RIGHT, typename R>
__device__ __forceinline__ void xform_inplace(const double* __restrict__ xf, R* V) {
  constexpr int NI = ncart(L), NO = AO<L, PURE>::N;
#pragma unroll
  for (int l = 0; l < LEFT; ++l)
#pragma unroll
    for (int r = 0; r < RIGHT; ++r) {
      R t[NI];
#pragma unroll
      for (int i = 0; i < NI; ++i) t[i] = V[(l * NI + i) * RIGHT + r];
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        R s = 0;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const double c = xf[(((L - 2) * 2 + (PURE ? 1 : 0)) * 10 + o) * 10 + i];
          s += (R)c * t[i];
        }
        V[(l * NO + o) * RIGHT + r] = s;
      }
    }
}

// Bra primitive pairs staged in LDS by a block whose bra is fixed (k_jkb):
// every lane reads pair i from the same LDS address (a broadcast), with no
// vector-memory access in the quadrature loop.  Up to JKB_MAX_BRA_PRIMS pairs.
#define JKB_MAX_BRA_PRIMS 128

// Contracted (e0|f0) of a quartet: Rys quadrature over every surviving
// primitive quartet.  The FP64 path reads the PrimJK records, the FP32 path
// the PrimF32 copies (prims32).  UNIFORM_BRA: the caller guarantees one bra per
// wavefront; its primitive pairs come from the LDS copy sbra when given, else
// through the constant address space (scalar cache).  imu2: 1/mu^2 of the
// erf-attenuated operator, 0 for 1/r.
template <int LA, int LB, int LC, int LD, typename R, bool UNIFORM_BRA = false>
__device__ __forceinline__ void contract(const ShellPair& P, const ShellPair& Q, const PrimJK* __restrict__ prims,
                                         const PrimF32* __restrict__ prims32, const double* __restrict__ tabN,
                                         const float* __restrict__ tab32N, R qcut, R imu2,
                                         R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF],
                                         int& nprim_used, const PrimV<double>* sbra = nullptr,
                                         const PrimV<float>* sbra32 = nullptr, int nbra64 = 0,
                                         int nbra32 = 0) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr bool F32 = sizeof(R) == 4;
#pragma unroll
  for (int e = 0; e < Dm::NE; ++e)
#pragma unroll
    for (int f = 0; f < Dm::NF; ++f) acc[e][f] = 0;
  R AC[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) AC[d] = (R)(P.A[d] - Q.A[d]);  // formed in FP64
  const R qcut2 = qcut * qcut;
  nprim_used = 0;
  auto ket = [&](int i) -> PrimV<R> {
    if constexpr (F32) return prim_v<float>(prims32[i]);
    else return prim_v<double>(prims[i]);
  };
  if (UNIFORM_BRA && sbra) {
    // the LDS copies hold only the bra primitive pairs with |K| >= qcut
    // (nbra64 FP64, nbra32 FP32 records), so the loop needs no test for them
    const int npp = __builtin_amdgcn_readfirstlane(sbra32 ? nbra32 : nbra64);
    // small classes, FP32: the next ket primitive pair is loaded while the
    // current one runs against the bra's primitives (where registers are
    // short, i.e. the larger classes and FP64, it costs more than it hides)
    constexpr bool PREF = F32 && Dm::NE * Dm::NF <= 9;
    const int nq = Q.nprim;
    PrimV<R> qn{};
    if (PREF && nq > 0) qn = ket(Q.prim0);
    for (int iq = 0; iq < nq; ++iq) {
      PrimV<R> qq;
      if constexpr (PREF) {
        qq = qn;
        if (iq + 1 < nq) qn = ket(Q.prim0 + iq + 1);
      } else {
        qq = ket(Q.prim0 + iq);
      }
      if (fabs(qq.Kg) < qcut) continue;
      R AQ[3];
      ket_offset<LA, LB, LC, LD, R>(AC, qq, AQ);
      auto run = [&](auto bra) {
        for (int ip = 0; ip < npp; ++ip)
          nprim_used += rys_accumulate<LA, LB, LC, LD, R>(bra(ip), qq, AQ, qcut2, imu2, tabN, tab32N, acc) ? 1 : 0;
      };
      if constexpr (F32) {
        if (sbra32) run([&](int i) { return sbra32[i]; });
        else run([&](int i) { return prim_v<float>(sbra[i]); });
      } else {
        run([&](int i) { return sbra[i]; });
      }
    }
    return;
  }
  const int pp0 = UNIFORM_BRA ? __builtin_amdgcn_readfirstlane(P.prim0) : P.prim0;
  const int npp = UNIFORM_BRA ? __builtin_amdgcn_readfirstlane(P.nprim) : P.nprim;
  auto bra = [&](int i) -> PrimV<R> {
    if constexpr (F32) return prim_v<float>(UNIFORM_BRA ? load_uniform(prims32, i) : prims32[i]);
    else return prim_v<double>(UNIFORM_BRA ? load_uniform(prims, i) : prims[i]);
  };
  for (int iq = 0; iq < Q.nprim; ++iq) {
    const PrimV<R> qq = ket(Q.prim0 + iq);
    if (fabs(qq.Kg) < qcut) continue;
    R AQ[3];
    ket_offset<LA, LB, LC, LD, R>(AC, qq, AQ);
    for (int ip = 0; ip < npp; ++ip) {
      const PrimV<R> pp = bra(pp0 + ip);
      if (fabs(pp.Kg) < qcut) continue;
      nprim_used += rys_accumulate<LA, LB, LC, LD, R>(pp, qq, AQ, qcut2, imu2, tabN, tab32N, acc) ? 1 : 0;
    }
  }
}

// Large classes: the AO block is produced one bra AO pair (a', b') at a time.
// The bra HRR and bra AO transform are evaluated on the fly from the
// contracted set, then the ket HRR and ket AO transform give the slice
// S[c'][d'], handed to vis(a', b', S).  Only acc and small slices stay live.
template <int LA, int LB, int LC, int LD, bool PURE, typename Vis>
__device__ __forceinline__ void stream_slices(const ShellPair& P, const ShellPair& Q, const double* __restrict__ xf,
                                              const double (&acc)[Dims<LA, LB, LC, LD>::NE]
                                                                 [Dims<LA, LB, LC, LD>::NF],
                                              Vis&& vis) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NA = Dm::NA, NB = Dm::NB, NC = Dm::NC, ND = Dm::ND, NF = Dm::NF;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N;
  double pa[3][LB + 1], pc[3][LD + 1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double ab = P.A[d] - P.B[d], cd = Q.A[d] - Q.B[d];
    pa[d][0] = 1.0;
    pc[d][0] = 1.0;
#pragma unroll
    for (int k = 1; k <= LB; ++k) pa[d][k] = pa[d][k - 1] * ab;
#pragma unroll
    for (int k = 1; k <= LD; ++k) pc[d][k] = pc[d][k - 1] * cd;
  }
  auto xa = [&](int L, int o, int i) -> double {  // AO row o, canonical col i
    if (L < 2) return o == i ? 1.0 : 0.0;
    return xf[(((L - 2) * 2 + (PURE ? 1 : 0)) * 10 + o) * 10 + i];
  };
#pragma unroll 1
  for (int ao = 0; ao < OA; ++ao)
#pragma unroll 1
    for (int bo = 0; bo < OB; ++bo) {
      double y[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) y[f] = 0.0;
      sfor<NA>([&](auto Ai) {
        constexpr int a = OQPH_CI(Ai);
        const double wa = xa(LA, ao, a);
        if (wa == 0.0) return;
        constexpr int ax = c_lx(LA, a), ay = c_ly(LA, a), az = c_lz(LA, a);
        sfor<NB>([&](auto Bi) {
          constexpr int b = OQPH_CI(Bi);
          const double wab = wa * xa(LB, bo, b);
          if (wab == 0.0) return;
          constexpr int bx = c_lx(LB, b), by = c_ly(LB, b), bz = c_lz(LB, b);
          sfor<bx + 1>([&](auto JX) {
            constexpr int jx = OQPH_CI(JX);
            sfor<by + 1>([&](auto JY) {
              constexpr int jy = OQPH_CI(JY);
              sfor<bz + 1>([&](auto JZ) {
                constexpr int jz = OQPH_CI(JZ);
                constexpr int e = range_index<LA>(LA + jx + jy + jz, c_idx(ax + jx, ay + jy, az + jz));
                constexpr int cb = binom(bx, jx) * binom(by, jy) * binom(bz, jz);
                const double cf = wab * (double)cb * pa[0][bx - jx] * pa[1][by - jy] * pa[2][bz - jz];
#pragma unroll
                for (int f = 0; f < NF; ++f) y[f] += cf * acc[e][f];
              });
            });
          });
        });
      });
      double sl[NC * ND];
      sfor<NC>([&](auto Ci) {
        constexpr int c = OQPH_CI(Ci);
        constexpr int cx = c_lx(LC, c), cy = c_ly(LC, c), cz = c_lz(LC, c);
        sfor<ND>([&](auto Di) {
          constexpr int dd = OQPH_CI(Di);
          constexpr int dx = c_lx(LD, dd), dy = c_ly(LD, dd), dz = c_lz(LD, dd);
          double v = 0.0;
          sfor<dx + 1>([&](auto JX) {
            constexpr int jx = OQPH_CI(JX);
            sfor<dy + 1>([&](auto JY) {
              constexpr int jy = OQPH_CI(JY);
              sfor<dz + 1>([&](auto JZ) {
                constexpr int jz = OQPH_CI(JZ);
                constexpr int f = range_index<LC>(LC + jx + jy + jz, c_idx(cx + jx, cy + jy, cz + jz));
                constexpr int cb = binom(dx, jx) * binom(dy, jy) * binom(dz, jz);
                v += (double)cb * pc[0][dx - jx] * pc[1][dy - jy] * pc[2][dz - jz] * y[f];
              });
            });
          });
          sl[c * ND + dd] = v;
        });
      });
      if constexpr (LC >= 2) xform_inplace<LC, PURE, 1, ND, double>(xf, sl);
      if constexpr (LD >= 2) xform_inplace<LD, PURE, AO<LC, PURE>::N, 1, double>(xf, sl);
      vis(ao, bo, sl);
    }
}

template <int LA, int LB, int LC, int LD>
struct Big {
  static constexpr bool value = Dims<LA, LB, LC, LD>::NBLK > 256;
};

// Final-representation block of a quartet: V[a][b][c][d] in AO order.
template <int LA, int LB, int LC, int LD, bool PURE, typename R, bool UNIFORM_BRA = false>
__device__ __forceinline__ void quartet_block(const ShellPair& P, const ShellPair& Q, const PrimJK* __restrict__ prims,
                                              const PrimF32* __restrict__ prims32, const double* __restrict__ tabN,
                                              const float* __restrict__ tab32N, const double* __restrict__ xf,
                                              R qcut, R imu2, R (&V)[Dims<LA, LB, LC, LD>::NBLK], int& nprim_used,
                                              const PrimV<double>* sbra = nullptr,
                                              const PrimV<float>* sbra32 = nullptr, int nbra64 = 0,
                                              int nbra32 = 0) {
  using Dm = Dims<LA, LB, LC, LD>;
  R acc[Dm::NE][Dm::NF];
  contract<LA, LB, LC, LD, R, UNIFORM_BRA>(P, Q, prims, prims32, tabN, tab32N, qcut, imu2, acc, nprim_used, sbra,
                                           sbra32, nbra64, nbra32);
  R AB[3], CD[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = (R)(P.A[d] - P.B[d]);
    CD[d] = (R)(Q.A[d] - Q.B[d]);
  }
  R Y[Dm::NA * Dm::NB][Dm::NF];
  hrr_bra<LA, LB, Dm::NF, R>(acc, AB, Y);
  hrr_ket<LC, LD, Dm::NA * Dm::NB, R>(Y, CD, V);
  // AO transform index by index, in place (only L >= 2 indices change); the
  // AO block ends up in the first OA*OB*OC*OD entries of V
  constexpr int NB = Dm::NB, NC = Dm::NC, ND = Dm::ND;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N;
  if constexpr (LA >= 2) xform_inplace<LA, PURE, 1, NB * NC * ND, R>(xf, V);
  if constexpr (LB >= 2) xform_inplace<LB, PURE, OA, NC * ND, R>(xf, V);
  if constexpr (LC >= 2) xform_inplace<LC, PURE, OA * OB, ND, R>(xf, V);
  if constexpr (LD >= 2) xform_inplace<LD, PURE, OA * OB * OC, 1, R>(xf, V);
}

__device__ __forceinline__ void fatomic(double* p, double v) { unsafeAtomicAdd(p, v); }

// one atomic per wavefront for a per-thread predicate (all active lanes call it)
__device__ __forceinline__ void sum_wave(unsigned long long* ctr, unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const unsigned long long act = __ballot(1);
  if (v != 0ull && __lane_id() == (unsigned)(__ffsll((long long)act) - 1)) atomicAdd(ctr, v);
}

// sum over all 64 lanes (call with the whole wavefront active)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ void count_wave(unsigned long long* ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  const unsigned long long act = __ballot(1);
  if (m != 0ull && __lane_id() == (unsigned)(__ffsll((long long)act) - 1))
    atomicAdd(ctr, (unsigned long long)__popcll(m));
}

// Root tables of the class's root count are copied into LDS once per block
// when they fit (FP64 for up to 3 roots, FP32 for every class that has an
// FP32 path): the table lookups are per-lane gathers at random intervals,
// which LDS serves at a fraction of the vector-memory cost.
template <int NR>
struct RysLds {
  static constexpr bool f64 = rys_len<NR>() * 8 <= 36 * 1024;
};

// TWO: the open-shell pass (prm.D2 / prm.F2 set), compiled separately so the
// closed-shell kernels keep their register allocation.
template <int LA, int LB, int LC, int LD, bool PURE, bool TWO>
__global__ __launch_bounds__(JK_BLOCK) void k_jk(JKParams prm, WaveWork ww) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NR = Dm::NROOT;
  constexpr bool T64 = RysLds<NR>::f64 && !Big<LA, LB, LC, LD>::value;
  constexpr bool T32 = Dm::NE * Dm::NF <= 256 && !Big<LA, LB, LC, LD>::value;  // classes with an FP32 path
  __shared__ double s_t64[T64 ? rys_len<NR>() : 1];
  __shared__ float s_t32[T32 ? rys32_len<NR>() : 1];
  if constexpr (T64)
    for (int i = threadIdx.x; i < rys_len<NR>(); i += blockDim.x) s_t64[i] = prm.rys[RYS_OFF[NR] + i];
  if constexpr (T32)
    if (prm.mixed)
      for (int i = threadIdx.x; i < rys32_len<NR>(); i += blockDim.x) s_t32[i] = prm.rys32[RYS32_OFF[NR] + i];
  __syncthreads();
  const double* tabN = T64 ? (const double*)s_t64 : prm.rys + RYS_OFF[NR];
  const float* tab32N = T32 ? (const float*)s_t32 : prm.rys32 + RYS32_OFF[NR];

  // one wavefront = one chunk of 64 kets of one segment; the bra is uniform
  const long long wave = __builtin_amdgcn_readfirstlane(
      (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6));
  if (wave >= ww.nwaves) return;  // wave-uniform
  int lo = 0, hi = ww.nseg;
  while (hi - lo > 1) {  // scalar binary search over the chunk prefix
    const int mid = (lo + hi) >> 1;
    if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
  }
  const int seg = __builtin_amdgcn_readfirstlane(lo);
  const int k = (int)(wave - ww.choff[seg]) * 64 + (int)__lane_id();
  const bool active = k < ww.seg_n[seg];
  const int ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
  int iq = 0;
  const ShellPair P = prm.pairs[ip];
  ShellPair Q{};
  bool skip = true;
  double bound = 0.0;
  bool mine = true;  // the quartet belongs to this share (OpenQP's MPI split, int2.F90:759-761)
  if (active) {
    iq = ww.ket[ww.seg_k0[seg] + k];
    Q = prm.pairs[iq];
    // ---- OpenQP screening (int2.F90:763-800)
    const bool p_first = (P.si > Q.si) || (P.si == Q.si && P.sj >= Q.sj);
    if (prm.own) mine = prm.own[p_first ? ip : iq] != 0;
    const double qbra = p_first ? P.Q : Q.Q;
    const int n = prm.nsh;
    const double* ds = prm.dsh;
    double dm = fmax(4.0 * ds[P.si * n + P.sj], 4.0 * ds[Q.si * n + Q.sj]);
    dm = fmax(dm, fmax(ds[P.sj * n + Q.sj], ds[P.sj * n + Q.si]));
    dm = fmax(dm, fmax(ds[P.si * n + Q.sj], ds[P.si * n + Q.si]));
    bound = P.Q * Q.Q * dm;
    skip = !mine || (qbra * prm.maxden < prm.cutoff) || (bound < prm.cutoff);
  }
  unsigned long long* ctr = prm.counters + 8 * (int)(wave & (OQPH_CTR_SLOTS - 1));
  // every lane of the wave reaches the counters (no early exit before them)
  count_wave(ctr + 1, active && mine && skip);
  count_wave(ctr + 0, !skip);
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N, OD = AO<LD, PURE>::N;
  if constexpr (Big<LA, LB, LC, LD>::value) {
    // ---- large classes: FP64, AO slices streamed into the digestion
    using Dm = Dims<LA, LB, LC, LD>;
    double acc[Dm::NE][Dm::NF];
    int nprim = 0;
    if (!skip)
      contract<LA, LB, LC, LD, double, true>(P, Q, prm.prims, prm.prims32, tabN, tab32N, 1.0e-4 * prm.cutoff,
                                             prm.inv_mu2, acc, nprim);
    count_wave(ctr + 2, false);
    sum_wave(ctr + 3, (unsigned long long)nprim * flops_prim<LA, LB, LC, LD>() +
                          (nprim ? (TWO ? 24ull : 12ull) * OA * OB * OC * OD : 0ull));
    sum_wave(ctr + 4, (unsigned long long)nprim);
    if (skip || nprim == 0) return;
    double deg = 1.0;
    if (P.sa == P.sb) deg *= 0.5;
    if (Q.sa == Q.sb) deg *= 0.5;
    if (ip == iq) deg *= 0.5;
    const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se, cut = prm.cutoff;
    const int oa = P.ao_a, ob = P.ao_b, oc = Q.ao_a, od = Q.ao_b;
    const double* Dp = prm.D;
    double* F = prm.F;
    double Dab[OA * OB], Dcd[OC * OD], Dac[OA * OC], Dad[OA * OD], Dbc[OB * OC], Dbd[OB * OD];
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int b = 0; b < OB; ++b) Dab[a * OB + b] = Dp[tri_idx(oa + a, ob + b)];
#pragma unroll
      for (int c = 0; c < OC; ++c) Dac[a * OC + c] = Dp[tri_idx(oa + a, oc + c)];
#pragma unroll
      for (int d = 0; d < OD; ++d) Dad[a * OD + d] = Dp[tri_idx(oa + a, od + d)];
    }
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int c = 0; c < OC; ++c) Dbc[b * OC + c] = Dp[tri_idx(ob + b, oc + c)];
#pragma unroll
      for (int d = 0; d < OD; ++d) Dbd[b * OD + d] = Dp[tri_idx(ob + b, od + d)];
    }
#pragma unroll
    for (int c = 0; c < OC; ++c)
#pragma unroll
      for (int d = 0; d < OD; ++d) Dcd[c * OD + d] = Dp[tri_idx(oc + c, od + d)];
    double Jcd[OC * OD], Kac[OA * OC], Kad[OA * OD], Kbc[OB * OC], Kbd[OB * OD];
#pragma unroll
    for (int i = 0; i < OC * OD; ++i) Jcd[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OA * OC; ++i) Kac[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OA * OD; ++i) Kad[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OB * OC; ++i) Kbc[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OB * OD; ++i) Kbd[i] = 0.0;
    stream_slices<LA, LB, LC, LD, PURE>(P, Q, prm.xform, acc, [&](int a, int b, double* S) {
      double jab = 0.0;
      const double dab = Dab[a * OB + b];
#pragma unroll
      for (int c = 0; c < OC; ++c)
#pragma unroll
        for (int d = 0; d < OD; ++d) {
          double v = S[c * OD + d];
          if (fabs(v) < cut) v = 0.0;  // OpenQP element cutoff on the AO value
          jab += v * Dcd[c * OD + d];
          Jcd[c * OD + d] += v * dab;
          Kac[a * OC + c] += v * Dbd[b * OD + d];
          Kad[a * OD + d] += v * Dbc[b * OC + c];
          Kbc[b * OC + c] += v * Dad[a * OD + d];
          Kbd[b * OD + d] += v * Dac[a * OC + c];
        }
      const int x = oa + a, y = ob + b;
      const double v = cj * jab * (x == y ? 2.0 : 1.0);
      if (v != 0.0) fatomic(F + tri_idx(x, y), v);
    });
    {  // J_cd into the ket's block of the ket-side Coulomb buffer
      double* Jd = prm.Jk + ww.jbase + (long long)(iq - ww.kofs) * (OC * OD);
#pragma unroll
      for (int c = 0; c < OC; ++c)
#pragma unroll
        for (int d = 0; d < OD; ++d) {
          const double v = cj * Jcd[c * OD + d];
          if (v != 0.0) fatomic(Jd + c * OD + d, v);
        }
    }
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int c = 0; c < OC; ++c) {
        const int x = oa + a, y = oc + c;
        const double v = ck * Kac[a * OC + c] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        const int x = oa + a, y = od + d;
        const double v = ck * Kad[a * OD + d] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
    }
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int c = 0; c < OC; ++c) {
        const int x = ob + b, y = oc + c;
        const double v = ck * Kbc[b * OC + c] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        const int x = ob + b, y = od + d;
        const double v = ck * Kbd[b * OD + d] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
    }
    if constexpr (!TWO) return;
    // open-shell pass: the exchange part for D2 into F2, a second sweep over
    // the AO slices of the same contracted integrals
    const double* D2 = prm.D2;
    double* F2 = prm.F2;
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int c = 0; c < OC; ++c) Dac[a * OC + c] = D2[tri_idx(oa + a, oc + c)];
#pragma unroll
      for (int d = 0; d < OD; ++d) Dad[a * OD + d] = D2[tri_idx(oa + a, od + d)];
    }
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int c = 0; c < OC; ++c) Dbc[b * OC + c] = D2[tri_idx(ob + b, oc + c)];
#pragma unroll
      for (int d = 0; d < OD; ++d) Dbd[b * OD + d] = D2[tri_idx(ob + b, od + d)];
    }
#pragma unroll
    for (int i = 0; i < OA * OC; ++i) Kac[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OA * OD; ++i) Kad[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OB * OC; ++i) Kbc[i] = 0.0;
#pragma unroll
    for (int i = 0; i < OB * OD; ++i) Kbd[i] = 0.0;
    stream_slices<LA, LB, LC, LD, PURE>(P, Q, prm.xform, acc, [&](int a, int b, double* S) {
#pragma unroll
      for (int c = 0; c < OC; ++c)
#pragma unroll
        for (int d = 0; d < OD; ++d) {
          double v = S[c * OD + d];
          if (fabs(v) < cut) v = 0.0;
          Kac[a * OC + c] += v * Dbd[b * OD + d];
          Kad[a * OD + d] += v * Dbc[b * OC + c];
          Kbc[b * OC + c] += v * Dad[a * OD + d];
          Kbd[b * OD + d] += v * Dac[a * OC + c];
        }
    });
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int c = 0; c < OC; ++c) {
        const int x = oa + a, y = oc + c;
        const double v = ck * Kac[a * OC + c] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F2 + tri_idx(x, y), v);
      }
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        const int x = oa + a, y = od + d;
        const double v = ck * Kad[a * OD + d] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F2 + tri_idx(x, y), v);
      }
    }
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int c = 0; c < OC; ++c) {
        const int x = ob + b, y = oc + c;
        const double v = ck * Kbc[b * OC + c] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F2 + tri_idx(x, y), v);
      }
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        const int x = ob + b, y = od + d;
        const double v = ck * Kbd[b * OD + d] * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F2 + tri_idx(x, y), v);
      }
    }
    return;
  } else {
  constexpr int NOUT = OA * OB * OC * OD;
  constexpr int NBLK = Dims<LA, LB, LC, LD>::NBLK;
  double V[NBLK];
#pragma unroll
  for (int i = 0; i < NOUT; ++i) V[i] = 0.0;
  int nprim = 0;
  // FP32 evaluation when every quartet of the wavefront has its bound below
  // the mixed-precision cutoff (a per-wavefront choice: no divergence
  // between the two paths); classes small enough to carry both paths
  bool lowp = false;
  if constexpr (T32) lowp = prm.mixed && __ballot(!skip && bound >= prm.fp32_cut) == 0ull;
  if (!skip) {
    if constexpr (T32) {
      if (lowp) {
        float Vf[NBLK];
        quartet_block<LA, LB, LC, LD, PURE, float, true>(P, Q, prm.prims, prm.prims32, tabN, tab32N, prm.xform,
                                                         (float)(1.0e-4 * prm.cutoff), (float)prm.inv_mu2, Vf,
                                                         nprim);
#pragma unroll
        for (int i = 0; i < NOUT; ++i) V[i] = (double)Vf[i];
      }
    }
    if (!lowp)
      quartet_block<LA, LB, LC, LD, PURE, double, true>(P, Q, prm.prims, prm.prims32, tabN, tab32N, prm.xform,
                                                        1.0e-4 * prm.cutoff, prm.inv_mu2, V, nprim);
  }
  count_wave(ctr + 2, lowp && !skip);
  sum_wave(ctr + 3, (unsigned long long)nprim * flops_prim<LA, LB, LC, LD>() +
                        (nprim ? (TWO ? 24ull : 12ull) * NOUT : 0ull));
  sum_wave(ctr + 4, (unsigned long long)nprim);
  const bool valid = !skip && nprim > 0;
  // all lanes stay converged through the digestion; invalid lanes carry V = 0
  const double c = prm.cutoff;
#pragma unroll
  for (int i = 0; i < NOUT; ++i)
    if (!valid || fabs(V[i]) < c) V[i] = 0.0;
  double deg = 1.0;
  if (P.sa == P.sb) deg *= 0.5;
  if (Q.sa == Q.sb) deg *= 0.5;
  if (ip == iq) deg *= 0.5;
  const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se;
  const int oa = P.ao_a, ob = P.ao_b, oc = Q.ao_a, od = Q.ao_b;
  const double* Dp = prm.D;
  double* F = prm.F;
  // wave-uniform bra: the J_ab block is reduced across the wavefront first
  const unsigned long long vmask = __ballot(valid);
  if (vmask == 0ull) return;  // wave-uniform exit: no lane has anything to digest
  const int lead = __ffsll((long long)vmask) - 1;  // first lane with a quartet to digest
  constexpr bool uni = true;  // the bra is uniform across the wavefront
  // J blocks
  {
    double Dcd[OC * OD], Dab[OA * OB];
#pragma unroll
    for (int x = 0; x < OC; ++x)
#pragma unroll
      for (int y = 0; y < OD; ++y) Dcd[x * OD + y] = Dp[tri_idx(oc + x, od + y)];
#pragma unroll
    for (int x = 0; x < OA; ++x)
#pragma unroll
      for (int y = 0; y < OB; ++y) Dab[x * OB + y] = Dp[tri_idx(oa + x, ob + y)];
#pragma unroll
    for (int a = 0; a < OA; ++a)
#pragma unroll
      for (int b = 0; b < OB; ++b) {
        double s = 0.0;
#pragma unroll
        for (int cd = 0; cd < OC * OD; ++cd) s += V[(a * OB + b) * OC * OD + cd] * Dcd[cd];
        const int x = oa + a, y = ob + b;
        const double v = cj * s * (x == y ? 2.0 : 1.0);
        if (uni) {
          const double t = wave_sum(v);
          // the leading valid lane owns the bra's AO offsets; invalid lanes may not
          if ((int)__lane_id() == lead && t != 0.0) fatomic(F + tri_idx(x, y), t);
        } else if (v != 0.0) {
          fatomic(F + tri_idx(x, y), v);
        }
      }
    // J_cd into the ket's block of the ket-side Coulomb buffer
    double* Jd = prm.Jk + ww.jbase + (long long)(iq - ww.kofs) * (OC * OD);
#pragma unroll
    for (int cc = 0; cc < OC; ++cc)
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        double s = 0.0;
#pragma unroll
        for (int ab = 0; ab < OA * OB; ++ab) s += V[ab * OC * OD + cc * OD + d] * Dab[ab];
        const double v = cj * s;
        if (v != 0.0) fatomic(Jd + cc * OD + d, v);
      }
  }
  // K blocks: K_ac += V D_bd, K_ad += V D_bc, K_bc += V D_ad, K_bd += V D_ac,
  // for D into F and, in the open-shell pass, for D2 into F2
#define VV(a, b, cc, d) V[(((a) * OB + (b)) * OC + (cc)) * OD + (d)]
  auto kblock = [&](const double* __restrict__ Dk, double* Fk) {
    double Dbd[OB * OD], Dbc[OB * OC], Dad[OA * OD], Dac[OA * OC];
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int d = 0; d < OD; ++d) Dbd[b * OD + d] = Dk[tri_idx(ob + b, od + d)];
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) Dbc[b * OC + cc] = Dk[tri_idx(ob + b, oc + cc)];
    }
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int d = 0; d < OD; ++d) Dad[a * OD + d] = Dk[tri_idx(oa + a, od + d)];
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) Dac[a * OC + cc] = Dk[tri_idx(oa + a, oc + cc)];
    }
#pragma unroll
    for (int a = 0; a < OA; ++a)
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < OB; ++b)
#pragma unroll
          for (int d = 0; d < OD; ++d) s += VV(a, b, cc, d) * Dbd[b * OD + d];
        const int x = oa + a, y = oc + cc;
        const double v = ck * s * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(Fk + tri_idx(x, y), v);
      }
#pragma unroll
    for (int a = 0; a < OA; ++a)
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < OB; ++b)
#pragma unroll
          for (int cc = 0; cc < OC; ++cc) s += VV(a, b, cc, d) * Dbc[b * OC + cc];
        const int x = oa + a, y = od + d;
        const double v = ck * s * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(Fk + tri_idx(x, y), v);
      }
#pragma unroll
    for (int b = 0; b < OB; ++b)
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < OA; ++a)
#pragma unroll
          for (int d = 0; d < OD; ++d) s += VV(a, b, cc, d) * Dad[a * OD + d];
        const int x = ob + b, y = oc + cc;
        const double v = ck * s * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(Fk + tri_idx(x, y), v);
      }
#pragma unroll
    for (int b = 0; b < OB; ++b)
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < OA; ++a)
#pragma unroll
          for (int cc = 0; cc < OC; ++cc) s += VV(a, b, cc, d) * Dac[a * OC + cc];
        const int x = ob + b, y = od + d;
        const double v = ck * s * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(Fk + tri_idx(x, y), v);
      }
  };
#undef VV
  kblock(Dp, F);
  if constexpr (TWO) kblock(prm.D2, prm.F2);
  }
}

// ---------------------------------------------------------------------------
// Bra-row kernel.  One block takes one bra pair (a b) with all of its
// surviving kets.  Every K update of a quartet -- K_ac, K_ad, K_bc, K_bd --
// lies in the AO rows of the bra shells a and b, and over the ~1e4 kets of one
// bra these rows are hit only at the AOs of the ket classes' shells.  The
// block therefore accumulates the rows in LDS (ds_add_f64) and flushes every
// nonzero entry with one global atomic at the end; J_ab stays in registers
// and is reduced once per wavefront; only J_cd goes to memory per quartet.
// Global FP64 atomics (about 2.4e10 lane-updates/s on the whole device) are
// what bounds the per-quartet digestion of k_jk; this cuts them 5-13x for the
// s/p/d classes.  Screening, quadrature and element cutoff are those of k_jk.
template <int LA, int LB, int LC, int LD>
struct JKB {
#ifndef OQPH_JKB_SMALL_THREADS
// block size of the classes with (e0|f0) sets of <= 4: 512 measured 1.3 %
// faster per build than 1024 and 6 % faster than 256 on (H2O)64/cc-pVDZ
#define OQPH_JKB_SMALL_THREADS 512
#endif
  // (e0|f0) sets of 5..9 ((ps|ps), (pp|ss), (ds|ss)): their blocks' K and
  // density rows span every p (or d) AO of the molecule, so one block fills
  // a CU's LDS; 1024 threads give that block 16 waves (the 128-VGPR cap of
  // that size costs these classes little: (pp|ss) 26.3 -> 21.2 ms, (ds|ss)
  // 22.3 -> 18.1 ms mixed on (H2O)64/cc-pVDZ, while the larger classes
  // spill heavily under it)
  static constexpr int NEF = Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF;
  static constexpr int threads = NEF <= 4 ? OQPH_JKB_SMALL_THREADS : (NEF <= 9 ? 1024 : 512);
  // (ss|ss) fits in 64 VGPRs without spilling: eight waves per SIMD;
  // OQPH_JKB_WAVES3 sets the target for (e0|f0) sets of 2..4 ((ps|ss))
#ifndef OQPH_JKB_WAVES3
#define OQPH_JKB_WAVES3 1
#endif
  static constexpr int min_waves = (Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF == 1)  ? 8
                                   : (Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF <= 4) ? OQPH_JKB_WAVES3
                                                                                                 : 1;
};

template <int LA, int LB, int LC, int LD, bool PURE, bool TWO>
__global__ __launch_bounds__((JKB<LA, LB, LC, LD>::threads))
__attribute__((amdgpu_waves_per_eu(JKB<LA, LB, LC, LD>::min_waves))) void k_jkb(JKParams prm, BraWork bw, ColMap cm) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NR = Dm::NROOT;
  constexpr bool T64 = RysLds<NR>::f64;
  constexpr bool T32 = Dm::NE * Dm::NF <= 256;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N, OD = AO<LD, PURE>::N;
  constexpr int NROW = OA + OB, NOUT = OA * OB * OC * OD, NBLK = Dm::NBLK;
  __shared__ double s_t64[T64 ? rys_len<NR>() : 1];
  __shared__ PrimV<double> s_bra[JKB_MAX_BRA_PRIMS];
  // FP32 copy of the bra primitives for the smallest classes, whose FP32
  // quadrature is short enough that converting the FP64 copy per primitive
  // quartet shows (5 % of (ps|ss)); larger classes keep the LDS for blocks
  constexpr bool B32 = T32 && Dm::NE * Dm::NF <= 4;
  __shared__ PrimV<float> s_bra32[B32 ? JKB_MAX_BRA_PRIMS : 1];
  __shared__ float s_t32[T32 ? rys32_len<NR>() : 1];
  __shared__ double s_jab[OA * OB];
  // dynamic LDS: K rows [NROW][ncol] (accumulated); when cm.drows, also D rows
  // [NROW][ncol] (the density at the same positions) and the bra shells' rows
  // of the shell-pair density maxima [2][nsh]; in the open-shell pass (F2
  // set) the K rows of D2 and, with cm.drows, the D2 rows follow
  extern __shared__ double s_dyn[];
  const int ncol = cm.ncol, nk = NROW * ncol, nsh = prm.nsh;
  const bool drows = cm.drows != 0;  // launch-uniform
  constexpr bool two = TWO;
  double* s_k = s_dyn;
  double* s_d = s_dyn + nk;
  double* s_ds = s_dyn + 2 * nk;
  double* s_k2 = drows ? s_ds + 2 * nsh : s_dyn + nk;
  double* s_d2 = s_k2 + nk;
  for (int i = threadIdx.x; i < nk; i += blockDim.x) s_k[i] = 0.0;
  if (two)
    for (int i = threadIdx.x; i < nk; i += blockDim.x) s_k2[i] = 0.0;
  if (threadIdx.x < OA * OB) s_jab[threadIdx.x] = 0.0;
  if constexpr (T64)
    for (int i = threadIdx.x; i < rys_len<NR>(); i += blockDim.x) s_t64[i] = prm.rys[RYS_OFF[NR] + i];
  if constexpr (T32)
    if (prm.mixed)
      for (int i = threadIdx.x; i < rys32_len<NR>(); i += blockDim.x) s_t32[i] = prm.rys32[RYS32_OFF[NR] + i];
  const int* item = bw.item + (size_t)blockIdx.x * (1 + OQPH_NBKT);
  const int ip = __builtin_amdgcn_readfirstlane(item[0]);
  const ShellPair P = prm.pairs[ip];
  const bool bra_lds = P.nprim <= JKB_MAX_BRA_PRIMS;  // block-uniform
  // LDS copies of the bra primitive pairs that pass the primitive screen
  // |K| >= 1e-4 c (FP64 test for the FP64 copy, FP32 test for the FP32 copy),
  // compacted by the first wavefront: the quadrature loops need no test
  __shared__ int s_nbra[2];
  if (bra_lds && threadIdx.x < 64) {
    const int lane0 = (int)__lane_id();
    const unsigned long long below = (1ull << lane0) - 1ull;
    const double qc64 = 1.0e-4 * prm.cutoff;
    const float qc32 = (float)qc64;
    int n64 = 0, n32 = 0;
    for (int i0 = 0; i0 < P.nprim; i0 += 64) {
      const int i = i0 + lane0;
      bool ok64 = false, ok32 = false;
      PrimV<double> r{};
      PrimV<float> r32{};
      if (i < P.nprim) {
        r = prim_v<double>(prm.prims[P.prim0 + i]);
        ok64 = !(fabs(r.Kg) < qc64);
        if (B32 && prm.mixed) {
          r32 = prim_v<float>(prm.prims32[P.prim0 + i]);
          ok32 = !(fabsf(r32.Kg) < qc32);
        }
      }
      const unsigned long long m64 = __ballot(ok64), m32 = __ballot(ok32);
      if (ok64) s_bra[n64 + __popcll(m64 & below)] = r;
      if constexpr (B32)
        if (ok32) s_bra32[n32 + __popcll(m32 & below)] = r32;
      n64 += __popcll(m64);
      n32 += __popcll(m32);
    }
    if (lane0 == 0) {
      s_nbra[0] = n64;
      s_nbra[1] = n32;
    }
  }
  // the density entries the digestion reads in the bra rows a and b, gathered
  // once per block instead of once per quartet
  if (drows) {
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const int r = i / ncol, col = i - r * ncol;
      const int x = r < OA ? P.ao_a + r : P.ao_b + (r - OA);
      const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
      s_d[i] = prm.D[tri_idx(x, y)];
      if (two) s_d2[i] = prm.D2[tri_idx(x, y)];
    }
    for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
      s_ds[i] = prm.dsh[P.si * nsh + i];
      s_ds[nsh + i] = prm.dsh[P.sj * nsh + i];
    }
  }
  __syncthreads();
  const double* tabN = T64 ? (const double*)s_t64 : prm.rys + RYS_OFF[NR];
  const float* tab32N = T32 ? (const float*)s_t32 : prm.rys32 + RYS32_OFF[NR];
  const PrimV<double>* sbra = bra_lds ? (const PrimV<double>*)s_bra : nullptr;
  const PrimV<float>* sbra32 = (B32 && bra_lds && prm.mixed) ? (const PrimV<float>*)s_bra32 : nullptr;
  const int nbra64 = bra_lds ? s_nbra[0] : 0, nbra32 = bra_lds ? s_nbra[1] : 0;
  const int lane = (int)__lane_id(), wv = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
  const int oa = P.ao_a, ob = P.ao_b;
  const double* Dp = prm.D;
  double* F = prm.F;
  double Dab[OA * OB];
#pragma unroll
  for (int x = 0; x < OA; ++x)
#pragma unroll
    for (int y = 0; y < OB; ++y) Dab[x * OB + y] = Dp[tri_idx(oa + x, ob + y)];
  double jab[OA * OB];
#pragma unroll
  for (int i = 0; i < OA * OB; ++i) jab[i] = 0.0;
  // per-lane statistics (32-bit; the model FLOPs follow from n_prim and n_valid)
  unsigned n_comp = 0, n_skip = 0, n_f32 = 0, n_valid = 0, n_prim = 0;

  for (int b = 0; b < OQPH_NBKT; ++b) {
    const int n = __builtin_amdgcn_readfirstlane(item[1 + b]);
    const int kbase = bw.kb[b];
    for (int k0 = wv * 64; k0 < n; k0 += nw * 64) {
      const int k = k0 + lane;
      const bool active = k < n;
      int iq = 0;
      ShellPair Q{};
      bool skip = true, mine = true;
      double bound = 0.0;
      if (active) {
        iq = bw.ket[kbase + k];
        Q = prm.pairs[iq];
        // ---- OpenQP screening (int2.F90:763-800); the bra shells' rows of
        // the shell-pair maxima come from LDS
        const bool p_first = (P.si > Q.si) || (P.si == Q.si && P.sj >= Q.sj);
        if (prm.own) mine = prm.own[p_first ? ip : iq] != 0;
        const double qbra = p_first ? P.Q : Q.Q;
        double dm = fmax(4.0 * prm.dsh[P.si * nsh + P.sj], 4.0 * prm.dsh[Q.si * nsh + Q.sj]);
        if (drows) {
          dm = fmax(dm, fmax(s_ds[nsh + Q.sj], s_ds[nsh + Q.si]));
          dm = fmax(dm, fmax(s_ds[Q.sj], s_ds[Q.si]));
        } else {
          const double* ds = prm.dsh;
          dm = fmax(dm, fmax(ds[P.sj * nsh + Q.sj], ds[P.sj * nsh + Q.si]));
          dm = fmax(dm, fmax(ds[P.si * nsh + Q.sj], ds[P.si * nsh + Q.si]));
        }
        bound = P.Q * Q.Q * dm;
        skip = !mine || (qbra * prm.maxden < prm.cutoff) || (bound < prm.cutoff);
      }
      n_skip += (active && mine && skip) ? 1 : 0;
      n_comp += skip ? 0 : 1;
      double V[NBLK];
#pragma unroll
      for (int i = 0; i < NOUT; ++i) V[i] = 0.0;
      int nprim = 0;
      bool lowp = false;
      if constexpr (T32) lowp = prm.mixed && __ballot(!skip && bound >= prm.fp32_cut) == 0ull;
      if (!skip) {
        if constexpr (T32) {
          if (lowp) {
            float Vf[NBLK];
            quartet_block<LA, LB, LC, LD, PURE, float, true>(P, Q, prm.prims, prm.prims32, tabN, tab32N,
                                                             prm.xform, (float)(1.0e-4 * prm.cutoff),
                                                             (float)prm.inv_mu2, Vf, nprim, sbra, sbra32, nbra64,
                                                             nbra32);
#pragma unroll
            for (int i = 0; i < NOUT; ++i) V[i] = (double)Vf[i];
          }
        }
        if (!lowp)
          quartet_block<LA, LB, LC, LD, PURE, double, true>(P, Q, prm.prims, prm.prims32, tabN, tab32N,
                                                            prm.xform, 1.0e-4 * prm.cutoff, prm.inv_mu2, V, nprim,
                                                            sbra, nullptr, nbra64, nbra32);
      }
      n_f32 += (lowp && !skip) ? 1 : 0;
      n_valid += nprim ? 1 : 0;
      n_prim += (unsigned)nprim;
      const bool valid = !skip && nprim > 0;
      if (__ballot(valid) == 0ull) continue;  // wave-uniform
      const double c = prm.cutoff;
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
        if (!valid || fabs(V[i]) < c) V[i] = 0.0;
      double deg = 1.0;
      if (P.sa == P.sb) deg *= 0.5;
      if (Q.sa == Q.sb) deg *= 0.5;
      if (ip == iq) deg *= 0.5;
      const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se;
      const int oc = Q.ao_a, od = Q.ao_b;
#define VV(a, b, cc, d) V[(((a) * OB + (b)) * OC + (cc)) * OD + (d)]
      // J_ab in registers; J_cd to memory
      {
        double Dcd[OC * OD];
#pragma unroll
        for (int x = 0; x < OC; ++x)
#pragma unroll
          for (int y = 0; y < OD; ++y) Dcd[x * OD + y] = valid ? Dp[tri_idx(oc + x, od + y)] : 0.0;
#pragma unroll
        for (int ab = 0; ab < OA * OB; ++ab) {
          double t = 0.0;
#pragma unroll
          for (int cd = 0; cd < OC * OD; ++cd) t += V[ab * OC * OD + cd] * Dcd[cd];
          jab[ab] += cj * t;
        }
        // J_cd into the ket's block of the ket-side Coulomb buffer
        double* Jd = prm.Jk + bw.jbase + (long long)(iq - bw.kofs) * (OC * OD);
#pragma unroll
        for (int cc = 0; cc < OC; ++cc)
#pragma unroll
          for (int d = 0; d < OD; ++d) {
            double t = 0.0;
#pragma unroll
            for (int ab = 0; ab < OA * OB; ++ab) t += V[ab * OC * OD + cc * OD + d] * Dab[ab];
            const double v = cj * t;
            if (v != 0.0) fatomic(Jd + cc * OD + d, v);
          }
      }
      // K rows a (0..OA-1) and b (OA..OA+OB-1) in LDS; the density in those
      // rows comes from the LDS copy at the same positions
      auto krows = [&](const double* sd, const double* __restrict__ Dg, double* sk) {
        const int colC = cm.colidx[Q.sa], colD = cm.dofs + cm.colidx[Q.sb];
        double Dbd[OB * OD], Dbc[OB * OC], Dad[OA * OD], Dac[OA * OC];
        if (drows) {
#pragma unroll
          for (int bb = 0; bb < OB; ++bb) {
#pragma unroll
            for (int d = 0; d < OD; ++d) Dbd[bb * OD + d] = sd[(OA + bb) * ncol + colD + d];
#pragma unroll
            for (int cc = 0; cc < OC; ++cc) Dbc[bb * OC + cc] = sd[(OA + bb) * ncol + colC + cc];
          }
#pragma unroll
          for (int a = 0; a < OA; ++a) {
#pragma unroll
            for (int d = 0; d < OD; ++d) Dad[a * OD + d] = sd[a * ncol + colD + d];
#pragma unroll
            for (int cc = 0; cc < OC; ++cc) Dac[a * OC + cc] = sd[a * ncol + colC + cc];
          }
        } else {
#pragma unroll
          for (int bb = 0; bb < OB; ++bb) {
#pragma unroll
            for (int d = 0; d < OD; ++d) Dbd[bb * OD + d] = Dg[tri_idx(ob + bb, od + d)];
#pragma unroll
            for (int cc = 0; cc < OC; ++cc) Dbc[bb * OC + cc] = Dg[tri_idx(ob + bb, oc + cc)];
          }
#pragma unroll
          for (int a = 0; a < OA; ++a) {
#pragma unroll
            for (int d = 0; d < OD; ++d) Dad[a * OD + d] = Dg[tri_idx(oa + a, od + d)];
#pragma unroll
            for (int cc = 0; cc < OC; ++cc) Dac[a * OC + cc] = Dg[tri_idx(oa + a, oc + cc)];
          }
        }
#pragma unroll
        for (int a = 0; a < OA; ++a) {
#pragma unroll
          for (int cc = 0; cc < OC; ++cc) {
            double t = 0.0;
#pragma unroll
            for (int bb = 0; bb < OB; ++bb)
#pragma unroll
              for (int d = 0; d < OD; ++d) t += VV(a, bb, cc, d) * Dbd[bb * OD + d];
            if (t != 0.0) atomicAdd(&sk[a * ncol + colC + cc], ck * t);
          }
#pragma unroll
          for (int d = 0; d < OD; ++d) {
            double t = 0.0;
#pragma unroll
            for (int bb = 0; bb < OB; ++bb)
#pragma unroll
              for (int cc = 0; cc < OC; ++cc) t += VV(a, bb, cc, d) * Dbc[bb * OC + cc];
            if (t != 0.0) atomicAdd(&sk[a * ncol + colD + d], ck * t);
          }
        }
#pragma unroll
        for (int bb = 0; bb < OB; ++bb) {
#pragma unroll
          for (int cc = 0; cc < OC; ++cc) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < OA; ++a)
#pragma unroll
              for (int d = 0; d < OD; ++d) t += VV(a, bb, cc, d) * Dad[a * OD + d];
            if (t != 0.0) atomicAdd(&sk[(OA + bb) * ncol + colC + cc], ck * t);
          }
#pragma unroll
          for (int d = 0; d < OD; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < OA; ++a)
#pragma unroll
              for (int cc = 0; cc < OC; ++cc) t += VV(a, bb, cc, d) * Dac[a * OC + cc];
            if (t != 0.0) atomicAdd(&sk[(OA + bb) * ncol + colD + d], ck * t);
          }
        }
      };
      if (valid) {
        krows(s_d, Dp, s_k);
        if constexpr (TWO) krows(s_d2, prm.D2, s_k2);
      }
#undef VV
    }
  }
  // J_ab: one LDS add per wavefront; statistics: one slot update per wavefront
#pragma unroll
  for (int ab = 0; ab < OA * OB; ++ab) {
    const double t = wave_sum(jab[ab]);
    if (lane == 0 && t != 0.0) atomicAdd(&s_jab[ab], t);
  }
  {
    unsigned long long* ctr = prm.counters + 8 * (int)((blockIdx.x * nw + wv) & (OQPH_CTR_SLOTS - 1));
    unsigned long long v[5] = {n_comp, n_skip, n_f32,
                               (unsigned long long)n_prim * flops_prim<LA, LB, LC, LD>() +
                                   (unsigned long long)n_valid * (two ? 24ull : 12ull) * NOUT,
                               n_prim};
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      unsigned long long t = v[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
      if (lane == 0 && t != 0ull) atomicAdd(ctr + i, t);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nk; i += blockDim.x) {
    const double v = s_k[i];
    if (v == 0.0) continue;
    const int r = i / ncol, col = i - r * ncol;
    const int x = r < OA ? oa + r : ob + (r - OA);
    const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
    fatomic(F + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
  }
  if (two)
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const double v = s_k2[i];
      if (v == 0.0) continue;
      const int r = i / ncol, col = i - r * ncol;
      const int x = r < OA ? oa + r : ob + (r - OA);
      const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
      fatomic(prm.F2 + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
    }
  if (threadIdx.x < OA * OB) {
    const double v = s_jab[threadIdx.x];
    const int x = oa + (int)threadIdx.x / OB, y = ob + (int)threadIdx.x % OB;
    if (v != 0.0) fatomic(F + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
  }
}

// Schwarz bounds Q_ab = sqrt(max |(ab|ab)|) for every pair of one class
// (ints_exchange, source/integrals/int2.F90:1582-1737); the pair records and
// primitive pairs passed here are built with the tight cutoffs of :1600-1604.
// imu2 = 1/mu^2 gives the bounds of the erf-attenuated integrals (the mu2
// argument of ints_exchange, :1597, :1625), 0 those of 1/r.
template <int LA, int LB, bool PURE>
__global__ __launch_bounds__(256) void k_schwarz(const ShellPair* pairs, const PrimJK* prims, const int* list,
                                                 int n, const double* tab, const double* xf,
                                                 double qcut, double imu2, double* Qout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int ip = list[t];
  const ShellPair P = pairs[ip];
  constexpr int N = AO<LA, PURE>::N * AO<LB, PURE>::N * AO<LA, PURE>::N * AO<LB, PURE>::N;
  double m = 0.0;
  int nprim;
  if constexpr (Big<LA, LB, LA, LB>::value) {
    double acc[Dims<LA, LB, LA, LB>::NE][Dims<LA, LB, LA, LB>::NF];
    contract<LA, LB, LA, LB, double>(P, P, prims, nullptr, tab + RYS_OFF[Dims<LA, LB, LA, LB>::NROOT], nullptr,
                                     qcut, imu2, acc, nprim);
    if (nprim)
      stream_slices<LA, LB, LA, LB, PURE>(P, P, xf, acc, [&](int, int, double* S) {
#pragma unroll
        for (int i = 0; i < AO<LA, PURE>::N * AO<LB, PURE>::N; ++i) m = fmax(m, fabs(S[i]));
      });
  } else {
    double V[Dims<LA, LB, LA, LB>::NBLK];
    quartet_block<LA, LB, LA, LB, PURE, double>(P, P, prims, nullptr, tab + RYS_OFF[Dims<LA, LB, LA, LB>::NROOT],
                                                nullptr, xf, qcut, imu2, V, nprim);
    if (nprim) {
#pragma unroll
      for (int i = 0; i < N; ++i) m = fmax(m, fabs(V[i]));
    }
  }
  Qout[ip] = sqrt(m);
}

}  // namespace oqph
