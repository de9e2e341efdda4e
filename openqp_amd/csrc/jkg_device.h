// jkg_device.h -- the grouped bra-row J/K kernel k_jkg, and its tiled form.
//
// A block takes one chunk: up to JKG::nm shell pairs of one class that share
// their two centres and every primitive pair's exponents (segmented shells
// contracted from one primitive set, e.g. oxygen 1s and 2s of cc-pVDZ over
// the same nine s exponents), with all of their surviving kets.  For each
// ket primitive pair and each chunk primitive pair the Rys quadrature runs
// once; each member adds the result times its own K/zeta, under its own
// primitive screen, so every member's contracted integrals equal those of the
// segmented evaluation (k_jkb) up to summation order.  The quartet screens
// of int2.F90:763-800 run per member and ket; the digestion, K rows in LDS
// (the members' distinct shells share rows) and J_ab in registers are those
// of k_jkb, and J_cd is summed over the members before its one atomic.
//
// Tiled form (TILE = true): the members are bra pairs of one class that share
// their first shell (its centre A), each with its own primitive pairs; the
// quadrature runs per member, and what the members share is the ket: its
// pair record and primitive pairs are read once for all of them, and its J_cd
// takes one atomic per tile instead of one per bra.
#pragma once
#include "jk_device.h"

namespace oqph {

// Members per chunk of a class: their (e0|f0) accumulators all stay in
// registers, so the cap falls with the size of the class's set.
template <int LA, int LB, int LC, int LD>
struct JKG {
  static constexpr int NEF = Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF;
#ifdef OQPH_JKG_NM
  static constexpr int nm = OQPH_JKG_NM;
#else
  // two shells of one family per side at most in the usual basis sets (the
  // generally contracted s pair of cc-pVXZ): 4 members for a bra of two s
  // shells, 2 otherwise
  static constexpr int nm = (LA == LB && NEF <= 4) ? 4 : (NEF <= 18 ? 2 : 1);
#endif
  static constexpr int threads = 512;
};

// Members per tile of the tiled form: the (e0|f0) accumulators and AO blocks
// of all members stay in registers.
template <int LA, int LB, int LC, int LD>
struct JKT {
  static constexpr int NEF = Dims<LA, LB, LC, LD>::NE * Dims<LA, LB, LC, LD>::NF;
  static constexpr int nm = NEF <= 4 ? 4 : (NEF <= 27 ? 2 : 1);
};

// Contracted (e0|f0) of every live member of a tile against one ket: bp
// holds the members' primitive pairs (LDS), member m's at [moff[m],
// moff[m + 1]); np[m] counts the primitive quartets its screens keep.
template <int LA, int LB, int LC, int LD, int NM, typename R>
__device__ __forceinline__ void contract_tile(const ShellPair& Q, const double (&Ac)[3],
                                              const PrimJK* __restrict__ prims, const PrimF32* __restrict__ prims32,
                                              const PrimV<R>* bp, const int* moff, int nmem, unsigned live,
                                              const double* __restrict__ tabN, const float* __restrict__ tab32N,
                                              R qcut, R imu2,
                                              R (&acc)[NM][Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF],
                                              int (&np)[NM]) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NE = Dm::NE, NF = Dm::NF;
  constexpr bool F32 = sizeof(R) == 4;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    np[m] = 0;
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int f = 0; f < NF; ++f) acc[m][e][f] = 0;
  }
  R AC[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) AC[d] = (R)(Ac[d] - Q.A[d]);  // formed in FP64
  const R qcut2 = qcut * qcut;
  auto ket = [&](int i) -> PrimV<R> {
    if constexpr (F32) return prim_v<float>(prims32[i]);
    else return prim_v<double>(prims[i]);
  };
  // small classes, FP32: the next ket primitive pair is loaded while the current one runs
  constexpr bool PREF = F32 && NE * NF <= 9;
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
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (m >= nmem) break;  // uniform
      if (!((live >> m) & 1u)) continue;
      const int p1 = moff[m + 1];
      for (int ip = moff[m]; ip < p1; ++ip) {
        const PrimV<R> pp = bp[ip];
        if (fabs(pp.Kg) < qcut) continue;
        np[m] += rys_accumulate<LA, LB, LC, LD, R>(pp, qq, AQ, qcut2, imu2, tabN, tab32N, acc[m]) ? 1 : 0;
      }
    }
  }
}

// Cartesian -> final AO block of one quartet from its contracted (e0|f0):
// horizontal transfer and AO transforms, as in quartet_block.
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
__device__ __forceinline__ void block_from_acc(const R (&AB)[3], const R (&CD)[3], const double* __restrict__ xf,
                                               const R (&acc)[Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF],
                                               R (&V)[Dims<LA, LB, LC, LD>::NBLK]) {
  using Dm = Dims<LA, LB, LC, LD>;
  R Y[Dm::NA * Dm::NB][Dm::NF];
  hrr_bra<LA, LB, Dm::NF, R>(acc, AB, Y);
  hrr_ket<LC, LD, Dm::NA * Dm::NB, R>(Y, CD, V);
  constexpr int NB = Dm::NB, NC = Dm::NC, ND = Dm::ND;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N;
  if constexpr (LA >= 2) xform_inplace<LA, PURE, 1, NB * NC * ND, R>(xf, V);
  if constexpr (LB >= 2) xform_inplace<LB, PURE, OA, NC * ND, R>(xf, V);
  if constexpr (LC >= 2) xform_inplace<LC, PURE, OA * OB, ND, R>(xf, V);
  if constexpr (LD >= 2) xform_inplace<LD, PURE, OA * OB * OC, 1, R>(xf, V);
}

// Contracted (e0|f0) of every live member against one ket.  rec: the
// chunk's primitive records in LDS (stride doubles or floats each); live:
// bit m set when member m's quartet passed the quartet screens.  np[m]
// counts the primitive quartets member m's own screens keep (the segmented
// count); the return value counts the chunk primitive quartets evaluated.
template <int LA, int LB, int LC, int LD, int NM, typename R>
__device__ __forceinline__ int contract_group(const ShellPair& Q, const double (&Ac)[3], const PrimJK* __restrict__ prims,
                                              const PrimF32* __restrict__ prims32, const R* __restrict__ rec,
                                              int ncp, int stride, int nmem, unsigned live,
                                              const double* __restrict__ tabN,
                                              const float* __restrict__ tab32N, R qcut, R imu2,
                                              R (&acc)[NM][Dims<LA, LB, LC, LD>::NE][Dims<LA, LB, LC, LD>::NF],
                                              int (&np)[NM]) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NE = Dm::NE, NF = Dm::NF;
  constexpr bool F32 = sizeof(R) == 4;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    np[m] = 0;
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int f = 0; f < NF; ++f) acc[m][e][f] = 0;
  }
  R AC[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) AC[d] = (R)(Ac[d] - Q.A[d]);  // formed in FP64
  const R qcut2 = qcut * qcut;
  int ng = 0;
  auto ket = [&](int i) -> PrimV<R> {
    if constexpr (F32) return prim_v<float>(prims32[i]);
    else return prim_v<double>(prims[i]);
  };
  // small classes, FP32: the next ket primitive pair is loaded while the current one runs
  constexpr bool PREF = F32 && NE * NF <= 9;
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
    const R kq = qq.Kg;
    if (fabs(kq) < qcut) continue;
    R AQ[3];
    ket_offset<LA, LB, LC, LD, R>(AC, qq, AQ);
    for (int ip = 0; ip < ncp; ++ip) {
      const R* r = rec + ip * stride;
      PrimV<R> pp;
      pp.g = r[0];
      pp.X[0] = r[1];
      pp.X[1] = r[2];
      pp.X[2] = r[3];
      pp.h = r[4];
      pp.Kg = r[5];
      const R zeta = pp.g, eta = qq.g;
      const R ze = zeta + eta + zeta * eta * imu2;
      const R fq = pp.Kg * kq;  // the largest member prefactor: no member passes below it
      if (fq * fq < qcut2 * ze) continue;
      const R sq = F32 ? (R)__builtin_amdgcn_rsqf((float)ze) : (R)rsq64((double)ze);
      const R ize = sq * sq;
      R val[NE][NF];
#pragma unroll
      for (int e = 0; e < NE; ++e)
#pragma unroll
        for (int f = 0; f < NF; ++f) val[e][f] = 0;
      rys_add<LA, LB, LC, LD, R>(pp, qq, AQ, ze, sq, ize, kq * sq, tabN, tab32N, val);
      ++ng;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        if (m >= nmem) break;  // uniform: records hold nmem prefactors
        // member m's own screens: its bra primitive prefactor and the
        // primitive-quartet prefactor (as the segmented contraction applies them)
        const R km = r[6 + m];
        const R t = km * kq;
        const bool pass = ((live >> m) & 1u) && fabs(km) >= qcut && t * t >= qcut2 * ze;
        const R w = pass ? km : (R)0;
        np[m] += pass ? 1 : 0;
#pragma unroll
        for (int e = 0; e < NE; ++e)
#pragma unroll
          for (int f = 0; f < NF; ++f) acc[m][e][f] += w * val[e][f];
      }
    }
  }
  return ng;
}

// TILE: the tiled form; its item[0] is the tile's offset in cv.pool (the
// launch's work lists) and cv.hdr, cv.cp, cv.cp32 are unused.
template <int LA, int LB, int LC, int LD, bool PURE, bool TWO, bool TILE>
__global__ __launch_bounds__(512) void k_jkg(JKParams prm, BraWork bw, ColMap cm, ChunkView cv) {
  using Dm = Dims<LA, LB, LC, LD>;
  constexpr int NM = TILE ? JKT<LA, LB, LC, LD>::nm : JKG<LA, LB, LC, LD>::nm;
  constexpr int NE = Dm::NE, NF = Dm::NF;
  constexpr int NR = Dm::NROOT;
  constexpr bool T64 = RysLds<NR>::f64;
  constexpr bool T32 = NE * NF <= 256;
  constexpr int OA = AO<LA, PURE>::N, OB = AO<LB, PURE>::N, OC = AO<LC, PURE>::N, OD = AO<LD, PURE>::N;
  constexpr int NOUT = OA * OB * OC * OD, NBLK = Dm::NBLK;
  constexpr bool SAME = (LA == LC && LB == LD);  // bra and ket pairs from one class list
  __shared__ double s_t64[T64 ? rys_len<NR>() : 1];
  __shared__ float s_t32[T32 ? rys32_len<NR>() : 1];
  __shared__ double s_jab[NM * OA * OB];
  __shared__ double s_dab[NM * OA * OB];  // each member's bra density block
  __shared__ double s_mq[NM], s_mthr[NM], s_mdab[NM];
  __shared__ int s_mi[NM][12];  // id, row a, row b, slot si, slot sj, si, sj, sa, sb, ao_a, ao_b
  __shared__ int s_moff[NM + 1];  // tiled form: each member's first primitive pair in the LDS copy
  __shared__ double s_mab[NM][3];  // tiled form: each member's A - B
  extern __shared__ double s_dyn[];
  const int* item = bw.item + (size_t)blockIdx.x * (1 + OQPH_NBKT);
  const int chunk = __builtin_amdgcn_readfirstlane(item[0]);
  const int* ch = TILE ? cv.pool + chunk : cv.pool + cv.hdr[chunk];
  const int nmem = __builtin_amdgcn_readfirstlane(ch[0]);
  const int cprim0 = __builtin_amdgcn_readfirstlane(ch[1]);
  // primitive records in LDS: the chunk's shared records (stride doubles
  // each), or the tile members' own primitive pairs (ch[2] of them, a
  // PrimV<double> and a PrimV<float> each)
  const int ncp = __builtin_amdgcn_readfirstlane(ch[2]);
  const int nrow = __builtin_amdgcn_readfirstlane(ch[3]);
  const int nshl = __builtin_amdgcn_readfirstlane(ch[4]);
  const int stride = TILE ? 6 : __builtin_amdgcn_readfirstlane(ch[5]);
  const int* rowao = ch + ch[6];
  const int* shl = ch + ch[7];
  const int ncol = cm.ncol, nk = nrow * ncol, nsh = prm.nsh;
  const bool drows = cm.drows != 0;  // launch-uniform
  constexpr bool two = TWO;  // the open-shell pass
  // dynamic LDS: K rows, D rows and dsh rows (drows), the K rows and D rows
  // of D2 (open-shell pass), the chunk's primitive records
  double* s_k = s_dyn;
  double* s_d = s_k + nk;
  double* s_ds = s_d + (drows ? nk : 0);
  double* s_k2 = s_ds + (drows ? (size_t)nshl * nsh : 0);
  double* s_d2 = s_k2 + (two ? nk : 0);
  double* s_cp = s_d2 + (two && drows ? nk : 0);
  float* s_cp32 = (float*)(s_cp + (size_t)ncp * stride);
  const int P0 = __builtin_amdgcn_readfirstlane(ch[OQPH_JKG_HDR]);
  const ShellPair PA = prm.pairs[P0];  // the members' common centres A, B
  for (int i = threadIdx.x; i < nk; i += blockDim.x) s_k[i] = 0.0;
  if (two)
    for (int i = threadIdx.x; i < nk; i += blockDim.x) s_k2[i] = 0.0;
  for (int i = threadIdx.x; i < NM * OA * OB; i += blockDim.x) s_jab[i] = 0.0;
  if constexpr (T64)
    for (int i = threadIdx.x; i < rys_len<NR>(); i += blockDim.x) s_t64[i] = prm.rys[RYS_OFF[NR] + i];
  if constexpr (T32)
    if (prm.mixed)
      for (int i = threadIdx.x; i < rys32_len<NR>(); i += blockDim.x) s_t32[i] = prm.rys32[RYS32_OFF[NR] + i];
  if (threadIdx.x < NM) {
    const int m = threadIdx.x;
    if (m < nmem) {
      const int* mi = ch + OQPH_JKG_HDR + OQPH_JKG_MEM * m;
      const ShellPair P = prm.pairs[mi[0]];
      s_mi[m][0] = mi[0];
      s_mi[m][1] = mi[1];
      s_mi[m][2] = mi[2];
      s_mi[m][3] = mi[3];
      s_mi[m][4] = mi[4];
      s_mi[m][5] = P.si;
      s_mi[m][6] = P.sj;
      s_mi[m][7] = P.sa;
      s_mi[m][8] = P.sb;
      s_mi[m][9] = P.ao_a;
      s_mi[m][10] = P.ao_b;
      s_mq[m] = P.Q;
      // the member's coarse ket list: Q_ket >= c / (4 maxD Q_bra), as the
      // host forms it (statistics of skipped quartets only)
      s_mthr[m] = prm.cutoff / (4.0 * prm.maxden * P.Q);
      s_mdab[m] = 4.0 * prm.dsh[P.si * nsh + P.sj];
#pragma unroll
      for (int d = 0; d < 3; ++d) s_mab[m][d] = P.A[d] - P.B[d];
    } else {
      s_mi[m][0] = -1;
      s_mq[m] = 0.0;
    }
  }
  if constexpr (TILE) {
    if (threadIdx.x == 0) {
      int o = 0;
      for (int m = 0; m < NM; ++m) {
        s_moff[m] = o;
        if (m < nmem) o += prm.pairs[ch[OQPH_JKG_HDR + OQPH_JKG_MEM * m]].nprim;
      }
      s_moff[NM] = o;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < NM * OA * OB; i += blockDim.x) {
    const int m = i / (OA * OB), ab = i - m * (OA * OB);
    double v = 0.0;
    if (m < nmem) {
      const int* mi = ch + OQPH_JKG_HDR + OQPH_JKG_MEM * m;
      const ShellPair& P = prm.pairs[mi[0]];
      v = prm.D[tri_idx(P.ao_a + ab / OB, P.ao_b + ab % OB)];
    }
    s_dab[i] = v;
  }
  if (drows) {
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const int r = i / ncol, col = i - r * ncol;
      const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
      s_d[i] = prm.D[tri_idx(rowao[r], y)];
      if (two) s_d2[i] = prm.D2[tri_idx(rowao[r], y)];
    }
    for (int i = threadIdx.x; i < nshl * nsh; i += blockDim.x) {
      const int sl = i / nsh, c = i - sl * nsh;
      s_ds[i] = prm.dsh[shl[sl] * nsh + c];
    }
  }
  if constexpr (TILE) {
    PrimV<double>* bp = (PrimV<double>*)s_cp;
    PrimV<float>* bp32 = (PrimV<float>*)s_cp32;
    for (int m = 0; m < nmem; ++m) {
      const ShellPair& P = prm.pairs[ch[OQPH_JKG_HDR + OQPH_JKG_MEM * m]];
      for (int i = threadIdx.x; i < P.nprim; i += blockDim.x) {
        bp[s_moff[m] + i] = prim_v<double>(prm.prims[P.prim0 + i]);
        if (prm.mixed) bp32[s_moff[m] + i] = prim_v<float>(prm.prims32[P.prim0 + i]);
      }
    }
  } else {
    for (int i = threadIdx.x; i < ncp * stride; i += blockDim.x) s_cp[i] = cv.cp[cprim0 + i];
    if (prm.mixed)
      for (int i = threadIdx.x; i < ncp * stride; i += blockDim.x) s_cp32[i] = cv.cp32[cprim0 + i];
  }
  __syncthreads();
  const double* tabN = T64 ? (const double*)s_t64 : prm.rys + RYS_OFF[NR];
  const float* tab32N = T32 ? (const float*)s_t32 : prm.rys32 + RYS32_OFF[NR];
  const int lane = (int)__lane_id(), wv = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
  const double* Dp = prm.D;
  double* F = prm.F;
  const double Ac[3] = {PA.A[0], PA.A[1], PA.A[2]};
  unsigned n_comp = 0, n_skip = 0, n_f32 = 0, n_valid = 0, n_prim = 0, n_grp = 0;
  const double c = prm.cutoff;

  for (int b = 0; b < OQPH_NBKT; ++b) {
    const int n = __builtin_amdgcn_readfirstlane(item[1 + b]);
    const int kbase = bw.kb[b];
    for (int k0 = wv * 64; k0 < n; k0 += nw * 64) {
      // the member tables are read from LDS where they are used: an opaque
      // zero offset keeps the compiler from hoisting every member's fields
      // out of the ket loop into registers
      int z;
      asm volatile("s_mov_b32 %0, 0" : "=s"(z));
      const int(*mi)[12] = (const int(*)[12])(&s_mi[0][0] + z);
      const double* mq = s_mq + z;
      const double* mthr = s_mthr + z;
      const double* mdab = s_mdab + z;
      const int k = k0 + lane;
      const bool active = k < n;
      int iq = 0;
      ShellPair Q{};
      unsigned live = 0;
      double bmax = 0.0;
      if (active) {
        iq = bw.ket[kbase + k];
        Q = prm.pairs[iq];
        const double dcd4 = 4.0 * prm.dsh[Q.si * nsh + Q.sj];
#pragma unroll 1
        for (int m = 0; m < nmem; ++m) {
          const int id = mi[m][0], si = mi[m][5], sj = mi[m][6];
          const bool tri = !SAME || iq <= id;  // each unordered pair of the class once
          // ---- OpenQP screening (int2.F90:763-800), per member
          const bool p_first = (si > Q.si) || (si == Q.si && sj >= Q.sj);
          const bool mine = prm.own ? prm.own[p_first ? id : iq] != 0 : true;
          const double qm = mq[m];
          const double qbra = p_first ? qm : Q.Q;
          double dm = fmax(mdab[m], dcd4);
          if (drows) {
            const double* ri = s_ds + mi[m][3] * nsh;
            const double* rj = s_ds + mi[m][4] * nsh;
            dm = fmax(dm, fmax(rj[Q.sj], rj[Q.si]));
            dm = fmax(dm, fmax(ri[Q.sj], ri[Q.si]));
          } else {
            const double* ds = prm.dsh;
            dm = fmax(dm, fmax(ds[sj * nsh + Q.sj], ds[sj * nsh + Q.si]));
            dm = fmax(dm, fmax(ds[si * nsh + Q.sj], ds[si * nsh + Q.si]));
          }
          const double bound = qm * Q.Q * dm;
          const bool skip = !tri || !mine || (qbra * prm.maxden < c) || (bound < c);
          if (!skip) {
            live |= 1u << m;
            bmax = fmax(bmax, bound);
          }
          n_skip += (tri && mine && skip && Q.Q >= mthr[m]) ? 1 : 0;
        }
      }
      n_comp += __popc(live);
      bool lowp = false;
      if constexpr (T32) lowp = prm.mixed && __ballot(live != 0u && bmax >= prm.fp32_cut) == 0ull;
      // each member's final AO block; the (e0|f0) accumulators live only
      // inside the precision branch that fills them
      double Vall[NM][NBLK];
      int np[NM];
#pragma unroll
      for (int m = 0; m < NM; ++m) np[m] = 0;
      double AB[3], CD[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        AB[d] = PA.A[d] - PA.B[d];
        CD[d] = Q.A[d] - Q.B[d];
      }
      if (live) {
        if constexpr (T32) {
          if (lowp) {
            float accf[NM][NE][NF];
            if constexpr (TILE)
              contract_tile<LA, LB, LC, LD, NM, float>(Q, Ac, prm.prims, prm.prims32, (const PrimV<float>*)s_cp32,
                                                       s_moff, nmem, live, tabN, tab32N, (float)(1.0e-4 * c),
                                                       (float)prm.inv_mu2, accf, np);
            else
              n_grp += contract_group<LA, LB, LC, LD, NM, float>(Q, Ac, prm.prims, prm.prims32, s_cp32, ncp, stride,
                                                                  nmem, live, tabN, tab32N, (float)(1.0e-4 * c),
                                                                  (float)prm.inv_mu2, accf, np);
            float ABf[3], CDf[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              ABf[d] = (float)AB[d];
              CDf[d] = (float)CD[d];
            }
#pragma unroll
            for (int m = 0; m < NM; ++m) {
              if (m >= nmem) break;  // uniform
              float Vf[NBLK];
              if constexpr (TILE)
#pragma unroll
                for (int d = 0; d < 3; ++d) ABf[d] = (float)s_mab[m][d];
              block_from_acc<LA, LB, LC, LD, PURE, float>(ABf, CDf, prm.xform, accf[m], Vf);
#pragma unroll
              for (int i = 0; i < NOUT; ++i) Vall[m][i] = (double)Vf[i];
            }
          }
        }
        if (!lowp) {
          double acc[NM][NE][NF];
          if constexpr (TILE)
            contract_tile<LA, LB, LC, LD, NM, double>(Q, Ac, prm.prims, prm.prims32, (const PrimV<double>*)s_cp,
                                                      s_moff, nmem, live, tabN, tab32N, 1.0e-4 * c, prm.inv_mu2, acc,
                                                      np);
          else
            n_grp += contract_group<LA, LB, LC, LD, NM, double>(Q, Ac, prm.prims, prm.prims32, s_cp, ncp, stride,
                                                                 nmem, live, tabN, tab32N, 1.0e-4 * c, prm.inv_mu2,
                                                                 acc, np);
#pragma unroll
          for (int m = 0; m < NM; ++m) {
            if (m >= nmem) break;  // uniform
            if constexpr (TILE)
#pragma unroll
              for (int d = 0; d < 3; ++d) AB[d] = s_mab[m][d];
            block_from_acc<LA, LB, LC, LD, PURE, double>(AB, CD, prm.xform, acc[m], Vall[m]);
          }
        }
      }
      n_f32 += lowp ? __popc(live) : 0;
      const int oc = Q.ao_a, od = Q.ao_b;
      bool any_valid = false;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        n_prim += (unsigned)np[m];
        n_valid += np[m] ? 1 : 0;
        any_valid |= ((live >> m) & 1u) && np[m] > 0;
      }
      if (__ballot(any_valid) == 0ull) continue;  // wave-uniform
      double Dcd[OC * OD];
#pragma unroll
      for (int x = 0; x < OC; ++x)
#pragma unroll
        for (int y = 0; y < OD; ++y) Dcd[x * OD + y] = any_valid ? Dp[tri_idx(oc + x, od + y)] : 0.0;
      double tcd[OC * OD];
#pragma unroll
      for (int i = 0; i < OC * OD; ++i) tcd[i] = 0.0;
      const int colC = active ? cm.colidx[Q.sa] : 0, colD = active ? cm.dofs + cm.colidx[Q.sb] : 0;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        if (m >= nmem) break;  // uniform
        const bool valid = ((live >> m) & 1u) && np[m] > 0;
        if (__ballot(valid) == 0ull) continue;  // wave-uniform
        double* V = Vall[m];
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
          if (!valid || fabs(V[i]) < c) V[i] = 0.0;  // OpenQP element cutoff
        const int id = mi[m][0];
        double deg = 1.0;
        if (mi[m][7] == mi[m][8]) deg *= 0.5;
        if (Q.sa == Q.sb) deg *= 0.5;
        if (id == iq) deg *= 0.5;
        const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se;
        const double* dab = s_dab + z + m * OA * OB;
#define VV(a, bb, cc, d) V[(((a) * OB + (bb)) * OC + (cc)) * OD + (d)]
        // J_ab of this member: reduced over the wavefront into LDS; J_cd
        // summed over the members
#pragma unroll
        for (int ab = 0; ab < OA * OB; ++ab) {
          double t = 0.0;
#pragma unroll
          for (int cd = 0; cd < OC * OD; ++cd) t += V[ab * OC * OD + cd] * Dcd[cd];
          t = wave_sum(cj * t);
          if (lane == 0 && t != 0.0) atomicAdd(&s_jab[m * OA * OB + ab], t);
        }
#pragma unroll
        for (int cd = 0; cd < OC * OD; ++cd) {
          double t = 0.0;
#pragma unroll
          for (int ab = 0; ab < OA * OB; ++ab) t += V[ab * OC * OD + cd] * dab[ab];
          tcd[cd] += cj * t;
        }
        // K rows of the member's shells a and b (LDS rows ra.., rb..)
        auto krows = [&](const double* sd, const double* __restrict__ Dg, double* sk) {
          const int ra = mi[m][1], rb = mi[m][2];
          const int oa = mi[m][9], ob = mi[m][10];
          double Dbd[OB * OD], Dbc[OB * OC], Dad[OA * OD], Dac[OA * OC];
          if (drows) {
#pragma unroll
            for (int bb = 0; bb < OB; ++bb) {
#pragma unroll
              for (int d = 0; d < OD; ++d) Dbd[bb * OD + d] = sd[(rb + bb) * ncol + colD + d];
#pragma unroll
              for (int cc = 0; cc < OC; ++cc) Dbc[bb * OC + cc] = sd[(rb + bb) * ncol + colC + cc];
            }
#pragma unroll
            for (int a = 0; a < OA; ++a) {
#pragma unroll
              for (int d = 0; d < OD; ++d) Dad[a * OD + d] = sd[(ra + a) * ncol + colD + d];
#pragma unroll
              for (int cc = 0; cc < OC; ++cc) Dac[a * OC + cc] = sd[(ra + a) * ncol + colC + cc];
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
              if (t != 0.0) atomicAdd(&sk[(ra + a) * ncol + colC + cc], ck * t);
            }
#pragma unroll
            for (int d = 0; d < OD; ++d) {
              double t = 0.0;
#pragma unroll
              for (int bb = 0; bb < OB; ++bb)
#pragma unroll
                for (int cc = 0; cc < OC; ++cc) t += VV(a, bb, cc, d) * Dbc[bb * OC + cc];
              if (t != 0.0) atomicAdd(&sk[(ra + a) * ncol + colD + d], ck * t);
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
              if (t != 0.0) atomicAdd(&sk[(rb + bb) * ncol + colC + cc], ck * t);
            }
#pragma unroll
            for (int d = 0; d < OD; ++d) {
              double t = 0.0;
#pragma unroll
              for (int a = 0; a < OA; ++a)
#pragma unroll
                for (int cc = 0; cc < OC; ++cc) t += VV(a, bb, cc, d) * Dac[a * OC + cc];
              if (t != 0.0) atomicAdd(&sk[(rb + bb) * ncol + colD + d], ck * t);
            }
          }
        };
        if (valid) {
          krows(s_d, Dp, s_k);
          if constexpr (TWO) krows(s_d2, prm.D2, s_k2);
        }
#undef VV
      }
      // J_cd of all members: one atomic per element, into the ket's block of
      // the ket-side Coulomb buffer
      double* Jd = prm.Jk + bw.jbase + (long long)(iq - bw.kofs) * (OC * OD);
#pragma unroll
      for (int cc = 0; cc < OC; ++cc)
#pragma unroll
        for (int d = 0; d < OD; ++d) {
          const double v = tcd[cc * OD + d];
          if (v != 0.0) fatomic(Jd + cc * OD + d, v);
        }
    }
  }
  // statistics: one slot update per wavefront
  {
    unsigned long long* ctr = prm.counters + 8 * (int)((blockIdx.x * nw + wv) & (OQPH_CTR_SLOTS - 1));
    unsigned long long v[6] = {n_comp, n_skip, n_f32,
                               (unsigned long long)n_prim * flops_prim<LA, LB, LC, LD>() +
                                   (unsigned long long)n_valid * (two ? 24ull : 12ull) * NOUT,
                               n_prim, TILE ? n_prim : n_grp};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
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
    const int x = rowao[r];
    const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
    fatomic(F + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
  }
  if (two)
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const double v = s_k2[i];
      if (v == 0.0) continue;
      const int r = i / ncol, col = i - r * ncol;
      const int x = rowao[r];
      const int y = col < cm.ncolC ? cm.aoC[col] : cm.aoD[col - cm.ncolC];
      fatomic(prm.F2 + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
    }
  for (int i = threadIdx.x; i < nmem * OA * OB; i += blockDim.x) {
    const double v = s_jab[i];
    if (v == 0.0) continue;
    const int m = i / (OA * OB), ab = i - m * (OA * OB);
    const int x = s_mi[m][9] + ab / OB, y = s_mi[m][10] + ab % OB;
    fatomic(F + tri_idx(x, y), v * (x == y ? 2.0 : 1.0));
  }
}

}  // namespace oqph
