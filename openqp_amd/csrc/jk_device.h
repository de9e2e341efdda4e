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

// apply the AO transform on index position IDX of a 4-index block
template <int L, bool PURE, int LEFT, int RIGHT, typename R>
__device__ __forceinline__ void xform_index(const double* __restrict__ xf, const R* in, R* out) {
  constexpr int NI = ncart(L), NO = AO<L, PURE>::N;
#pragma unroll
  for (int l = 0; l < LEFT; ++l)
#pragma unroll
    for (int o = 0; o < NO; ++o)
#pragma unroll
      for (int r = 0; r < RIGHT; ++r) {
        R s = 0;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const double t = xf[(((L - 2) * 2 + (PURE ? 1 : 0)) * 10 + o) * 10 + i];
          s += (R)t * in[(l * NI + i) * RIGHT + r];
        }
        out[(l * NO + o) * RIGHT + r] = s;
      }
}

// Final-representation block of a quartet: V[a][b][c][d] in AO order.
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
__device__ __forceinline__ void quartet_block(const ShellPair& P, const ShellPair& Q, const PrimPair* __restrict__ prims,
                                              const double* __restrict__ tab, const float* __restrict__ tab32,
                                              const double* __restrict__ herm,
                                              const double* __restrict__ xf, R qcut, R (&Vout)[AO<LA, PURE>::N * AO<LB, PURE>::N *
                                                                AO<LC, PURE>::N * AO<LD, PURE>::N],
                                              int& nprim_used) {
  using Dm = Dims<LA, LB, LC, LD>;
  R acc[Dm::NE][Dm::NF];
#pragma unroll
  for (int e = 0; e < Dm::NE; ++e)
#pragma unroll
    for (int f = 0; f < Dm::NF; ++f) acc[e][f] = 0;
  const R qcut2 = qcut * qcut;
  nprim_used = 0;
  for (int iq = 0; iq < Q.nprim; ++iq) {
    const PrimPair qq = prims[Q.prim0 + iq];
    if (fabs(qq.Kg) < (double)qcut) continue;
    for (int ip = 0; ip < P.nprim; ++ip) {
      const PrimPair pp = prims[P.prim0 + ip];
      if (fabs(pp.Kg) < (double)qcut) continue;
      nprim_used += rys_accumulate<LA, LB, LC, LD, R>(P.A, Q.A, pp, qq, qcut2, tab, tab32, herm, acc) ? 1 : 0;
    }
  }
  R AB[3], CD[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = (R)(P.A[d] - P.B[d]);
    CD[d] = (R)(Q.A[d] - Q.B[d]);
  }
  R Y[Dm::NA * Dm::NB][Dm::NF];
  hrr_bra<LA, LB, Dm::NF, R>(acc, AB, Y);
  R V[Dm::NBLK];
  hrr_ket<LC, LD, Dm::NA * Dm::NB, R>(Y, CD, V);
  // AO transform index by index (only L >= 2 indices change)
  constexpr int NA = Dm::NA, NB = Dm::NB, NC = Dm::NC, ND = Dm::ND;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N, OD = AO<LD, PURE>::N;
  if constexpr (LA < 2 && LB < 2 && LC < 2 && LD < 2) {
#pragma unroll
    for (int i = 0; i < Dm::NBLK; ++i) Vout[i] = V[i];
  } else {
    R W1[OA * NB * NC * ND];
    if constexpr (LA >= 2) xform_index<LA, PURE, 1, NB * NC * ND, R>(xf, V, W1);
    else {
#pragma unroll
      for (int i = 0; i < NA * NB * NC * ND; ++i) W1[i] = V[i];
    }
    R W2[OA * OB * NC * ND];
    if constexpr (LB >= 2) xform_index<LB, PURE, OA, NC * ND, R>(xf, W1, W2);
    else {
#pragma unroll
      for (int i = 0; i < OA * NB * NC * ND; ++i) W2[i] = W1[i];
    }
    R W3[OA * OB * OC * ND];
    if constexpr (LC >= 2) xform_index<LC, PURE, OA * OB, ND, R>(xf, W2, W3);
    else {
#pragma unroll
      for (int i = 0; i < OA * OB * NC * ND; ++i) W3[i] = W2[i];
    }
    if constexpr (LD >= 2) xform_index<LD, PURE, OA * OB * OC, 1, R>(xf, W3, Vout);
    else {
#pragma unroll
      for (int i = 0; i < OA * OB * OC * OD; ++i) Vout[i] = W3[i];
    }
  }
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

template <int LA, int LB, int LC, int LD, bool PURE>
__global__ __launch_bounds__(256) void k_jk(JKParams prm, ClassWork cw) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = w < cw.nwork;
  int ip = -1, iq = 0;
  ShellPair P{}, Q{};
  bool skip = true;
  double bound = 0.0;
  if (active) {
    // segment = bra * nbkt + ket bucket; the segment's kets are a prefix of the bucket
    int lo = 0, hi = cw.nbra * cw.nbkt;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cw.off[mid] <= w) lo = mid; else hi = mid;
    }
    ip = cw.bra[lo / cw.nbkt];
    iq = cw.ket[cw.kbs[lo % cw.nbkt] + (int)(w - cw.off[lo])];
    P = prm.pairs[ip];
    Q = prm.pairs[iq];
    // ---- OpenQP screening (int2.F90:763-800)
    const bool p_first = (P.si > Q.si) || (P.si == Q.si && P.sj >= Q.sj);
    const double qbra = p_first ? P.Q : Q.Q;
    const int n = prm.nsh;
    const double* ds = prm.dsh;
    double dm = fmax(4.0 * ds[P.si * n + P.sj], 4.0 * ds[Q.si * n + Q.sj]);
    dm = fmax(dm, fmax(ds[P.sj * n + Q.sj], ds[P.sj * n + Q.si]));
    dm = fmax(dm, fmax(ds[P.si * n + Q.sj], ds[P.si * n + Q.si]));
    bound = P.Q * Q.Q * dm;
    skip = (qbra * prm.maxden < prm.cutoff) || (bound < prm.cutoff);
  }
  unsigned long long* ctr = prm.counters;
  // every lane of the wave reaches the counters (no early exit before them)
  count_wave(ctr + 1, active && skip);
  count_wave(ctr + 0, !skip);
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N, OD = AO<LD, PURE>::N;
  constexpr int NOUT = OA * OB * OC * OD;
  double V[NOUT];
#pragma unroll
  for (int i = 0; i < NOUT; ++i) V[i] = 0.0;
  int nprim = 0;
  bool lowp = false;
  if (!skip) {
    // FP32 evaluation for quartets whose bound is below the mixed-precision
    // cutoff (classes small enough to carry both instantiations)
    if constexpr (Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF <= 256) {
      lowp = prm.mixed && bound < prm.fp32_cut;
      if (lowp) {
        float Vf[NOUT];
        quartet_block<LA, LB, LC, LD, PURE, float>(P, Q, prm.prims, prm.rys, prm.rys32, prm.herm, prm.xform,
                                                   (float)(1.0e-4 * prm.cutoff), Vf, nprim);
#pragma unroll
        for (int i = 0; i < NOUT; ++i) V[i] = (double)Vf[i];
      }
    }
    if (!lowp)
      quartet_block<LA, LB, LC, LD, PURE, double>(P, Q, prm.prims, prm.rys, prm.rys32, prm.herm, prm.xform,
                                                  1.0e-4 * prm.cutoff, V, nprim);
  }
  count_wave(ctr + 2, lowp);
  sum_wave(ctr + 3, (unsigned long long)nprim * flops_prim<LA, LB, LC, LD>() + (nprim ? 12ull * NOUT : 0ull));
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
  const int ip0 = __builtin_amdgcn_readlane(ip, lead);
  const bool uni = __ballot(ip == ip0 || !valid) == __ballot(1);
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
#pragma unroll
    for (int cc = 0; cc < OC; ++cc)
#pragma unroll
      for (int d = 0; d < OD; ++d) {
        double s = 0.0;
#pragma unroll
        for (int ab = 0; ab < OA * OB; ++ab) s += V[ab * OC * OD + cc * OD + d] * Dab[ab];
        const int x = oc + cc, y = od + d;
        const double v = cj * s * (x == y ? 2.0 : 1.0);
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
  }
  // K blocks: K_ac += V D_bd, K_ad += V D_bc, K_bc += V D_ad, K_bd += V D_ac
  {
    double Dbd[OB * OD], Dbc[OB * OC], Dad[OA * OD], Dac[OA * OC];
#pragma unroll
    for (int b = 0; b < OB; ++b) {
#pragma unroll
      for (int d = 0; d < OD; ++d) Dbd[b * OD + d] = Dp[tri_idx(ob + b, od + d)];
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) Dbc[b * OC + cc] = Dp[tri_idx(ob + b, oc + cc)];
    }
#pragma unroll
    for (int a = 0; a < OA; ++a) {
#pragma unroll
      for (int d = 0; d < OD; ++d) Dad[a * OD + d] = Dp[tri_idx(oa + a, od + d)];
#pragma unroll
      for (int cc = 0; cc < OC; ++cc) Dac[a * OC + cc] = Dp[tri_idx(oa + a, oc + cc)];
    }
#define VV(a, b, cc, d) V[(((a) * OB + (b)) * OC + (cc)) * OD + (d)]
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
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
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
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
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
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
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
        if (v != 0.0) fatomic(F + tri_idx(x, y), v);
      }
#undef VV
  }
}

// Schwarz bounds Q_ab = sqrt(max |(ab|ab)|) for every pair of one class
// (ints_exchange, source/integrals/int2.F90:1582-1737); the pair records and
// primitive pairs passed here are built with the tight cutoffs of :1600-1604.
template <int LA, int LB, bool PURE>
__global__ __launch_bounds__(256) void k_schwarz(const ShellPair* pairs, const PrimPair* prims, const int* list,
                                                 int n, const double* tab, const double* herm, const double* xf,
                                                 double qcut, double* Qout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int ip = list[t];
  const ShellPair P = pairs[ip];
  constexpr int N = AO<LA, PURE>::N * AO<LB, PURE>::N * AO<LA, PURE>::N * AO<LB, PURE>::N;
  double V[N];
  int nprim;
  quartet_block<LA, LB, LA, LB, PURE, double>(P, P, prims, tab, nullptr, herm, xf, qcut, V, nprim);
  double m = 0.0;
  if (nprim) {
#pragma unroll
    for (int i = 0; i < N; ++i) m = fmax(m, fabs(V[i]));
  }
  Qout[ip] = sqrt(m);
}

}  // namespace oqph
