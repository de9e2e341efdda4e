// hl_device.h -- shell quartets of the classes that have no class-templated
// kernel: the larger classes with an f shell (up to (ff|ff)), and the d
// classes whose templates would spill more than 8 KB per lane ((dd|dd) J/K,
// (dd|dd) and (dd|dp) gradient).  One wavefront evaluates one quartet at a
// time with every intermediate in LDS, sized per class at launch (HLDims).
// Covers the Fock build, the Schwarz bounds of the f pair classes and the 2e
// gradient for those classes; the algorithm is the one of the register
// kernels (Rys quadrature, contracted (e0|f0), binomial HRR, AO transform,
// OpenQP element cutoff and digestion), with loop bounds read at run time
// instead of unrolled, so (ff|ff) needs no register file larger than the
// hardware's.  Component exponents come from per-class tables and the HRR
// binomial-power coefficients from per-quartet tables, both in LDS.
//
// Reference: int2_rys_compute (source/integrals/int_rys.F90:156-785) for the
// integrals, int2.F90:763-800 / :1806-1812 for the screens and element cutoff,
// int2_rhf_data_t_update (int2.F90:1414-1484) for the digestion,
// ints_exchange (int2.F90:1582-1737) for the bounds, and grd2.F90:303-351 /
// grd2_rys.F90:303-937 for the gradient.
#pragma once
#include "jk_device.h"
#include "grad_params.h"

namespace oqph {

#define HL_BLOCK 64  // one wavefront per block
// occupancy hint of the run-time-class kernels (waves per SIMD): 3 caps them
// at 168 VGPRs with a few bytes of spill, 2.6 % faster than the unhinted 206
#ifndef OQPH_HL_WAVES
#define OQPH_HL_WAVES 3
#endif

// ---------------------------------------------------------- small helpers
__device__ __forceinline__ double hl_pow(double v, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= v;
  return r;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
// index of (x, y, z) inside the level range that starts at L0
__device__ __forceinline__ int hl_eidx(int L0, int x, int y, int z) {
  return ncum(x + y + z - 1) - ncum(L0 - 1) + c_idx(x, y, z);
}
// one node/weight pair of the nr-root quadrature, nr known at run time
__device__ __forceinline__ void hl_root(int nr, double T, const double* __restrict__ rys, int r, double& x,
                                        double& w) {
  switch (nr) {
    case 1: rys_root_one<1, double>(T, rys + RYS_OFF[1], nullptr, r, x, w); break;
    case 2: rys_root_one<2, double>(T, rys + RYS_OFF[2], nullptr, r, x, w); break;
    case 3: rys_root_one<3, double>(T, rys + RYS_OFF[3], nullptr, r, x, w); break;
    case 4: rys_root_one<4, double>(T, rys + RYS_OFF[4], nullptr, r, x, w); break;
    case 5: rys_root_one<5, double>(T, rys + RYS_OFF[5], nullptr, r, x, w); break;
    case 6: rys_root_one<6, double>(T, rys + RYS_OFF[6], nullptr, r, x, w); break;
    case 7: rys_root_one<7, double>(T, rys + RYS_OFF[7], nullptr, r, x, w); break;
    default: rys_root_one<8, double>(T, rys + RYS_OFF[8], nullptr, r, x, w); break;
  }
}

// Two-level index (i, j) of a flat index stepped by the wavefront width:
// one division at the start, then additions (j runs fastest, j < n2).
struct HLStep {
  int i, j, q, r, n2;
  __device__ __forceinline__ HLStep(int start, int n) : n2(n) {
    i = start / n;
    j = start - i * n;
    q = HL_BLOCK / n;
    r = HL_BLOCK - q * n;
  }
  __device__ __forceinline__ void next() {
    i += q;
    j += r;
    if (j >= n2) {
      j -= n2;
      ++i;
    }
  }
};

// Per-class tables (ints, LDS), filled once per block:
//   eT[ne], fT[nf]     exponents of the (e0| and |f0) level ranges, x | y<<4 | z<<8
//   abT[na nb], cdT[nc nd]  a | b<<4 | ax<<8 | ay<<12 | az<<16 | bx<<20 | by<<24 | bz<<28
//   tT[nt]             gradient transferred-table entries ia | ib<<4 | ic<<8 | id<<12
struct HLTabs {
  const int *eT, *fT, *abT, *cdT, *tT;
};
__device__ __forceinline__ HLTabs hl_tables(const HLDims& h, double* lds) {
  int* b = (int*)(lds + h.offI);
  HLTabs t;
  int* eT = b;
  int* fT = eT + h.ne;
  int* abT = fT + h.nf;
  int* cdT = abT + h.na * h.nb;
  int* tT = cdT + h.nc * h.nd;
  const int lane = threadIdx.x;
  for (int e = lane; e < h.ne; e += HL_BLOCK) {
    int L = h.la, c = e;
    while (c >= ncart(L)) { c -= ncart(L); ++L; }
    eT[e] = c_lx(L, c) | (c_ly(L, c) << 4) | (c_lz(L, c) << 8);
  }
  for (int f = lane; f < h.nf; f += HL_BLOCK) {
    int L = h.lc, c = f;
    while (c >= ncart(L)) { c -= ncart(L); ++L; }
    fT[f] = c_lx(L, c) | (c_ly(L, c) << 4) | (c_lz(L, c) << 8);
  }
  for (int t = lane; t < h.na * h.nb; t += HL_BLOCK) {
    const int a = t / h.nb, bb = t - a * h.nb;
    abT[t] = a | (bb << 4) | (c_lx(h.la, a) << 8) | (c_ly(h.la, a) << 12) | (c_lz(h.la, a) << 16) |
             (c_lx(h.lb, bb) << 20) | (c_ly(h.lb, bb) << 24) | (c_lz(h.lb, bb) << 28);
  }
  for (int t = lane; t < h.nc * h.nd; t += HL_BLOCK) {
    const int c = t / h.nd, d = t - c * h.nd;
    cdT[t] = c | (d << 4) | (c_lx(h.lc, c) << 8) | (c_ly(h.lc, c) << 12) | (c_lz(h.lc, c) << 16) |
             (c_lx(h.ld, d) << 20) | (c_ly(h.ld, d) << 24) | (c_lz(h.ld, d) << 28);
  }
  const int NTb = h.lb + 2, NTc = h.lc + 2, NTd = h.ld + 1;
  for (int t = lane; t < h.nt; t += HL_BLOCK) {
    const int ia = t / (NTb * NTc * NTd), r1 = t - ia * NTb * NTc * NTd, ib = r1 / (NTc * NTd),
              r2 = r1 - ib * NTc * NTd, ic = r2 / NTd, id = r2 - ic * NTd;
    tT[t] = ia | (ib << 4) | (ic << 8) | (id << 12);
  }
  __syncthreads();
  t.eT = eT; t.fT = fT; t.abT = abT; t.cdT = cdT; t.tT = tT;
  return t;
}

// HRR coefficients of one quartet (LDS, 123 doubles):
//   cab[x][b][j] = C(b, j) AB_x^(b-j), b <= 4;  ccd[x][d][m] = C(d, m) CD_x^(d-m), d <= 3
__device__ __forceinline__ void hl_coef(const ShellPair& P, const ShellPair& Q, double* C) {
  for (int t = threadIdx.x; t < 123; t += HL_BLOCK) {  // 123 entries over the 64 lanes
    if (t < 75) {
      const int x = t / 25, b = (t - 25 * x) / 5, j = t - 25 * x - 5 * b;
      const double v = x == 0 ? P.A[0] - P.B[0] : (x == 1 ? P.A[1] - P.B[1] : P.A[2] - P.B[2]);
      C[t] = j <= b ? (double)binom(b, j) * hl_pow(v, b - j) : 0.0;
    } else {
      const int u = t - 75, x = u / 16, d = (u - 16 * x) / 4, m = u - 16 * x - 4 * d;
      const double v = x == 0 ? Q.A[0] - Q.B[0] : (x == 1 ? Q.A[1] - Q.B[1] : Q.A[2] - Q.B[2]);
      C[t] = m <= d ? (double)binom(d, m) * hl_pow(v, d - m) : 0.0;
    }
  }
}

// Scalars of one primitive quartet shared by every lane (scalar members, so
// no lane indexes a private array with its run-time axis).
struct HLPrim {
  double PQx, PQy, PQz, PAx, PAy, PAz, QCx, QCy, QCz;
  double T, pref, rz, re, h2z, h2e, h2ze;
};
// J/K records (PrimJK): product centres relative to the pair's first centre,
// PQ = (A - C) + X_p - X_q; imu2 = 1/mu^2 of the erf-attenuated operator
// (zeta + eta -> zeta + eta + zeta eta / mu^2, as in rys_accumulate), 0 for 1/r.
__device__ __forceinline__ HLPrim hl_prim(const ShellPair& P, const ShellPair& Q, const PrimJK& pp, const PrimJK& qq,
                                          double imu2) {
  HLPrim s;
  const double zeta = pp.g, eta = qq.g, ze = zeta + eta + zeta * eta * imu2;
  const double sq = rsq64(ze), ize = sq * sq, rho = zeta * eta * ize;
  s.PQx = (P.A[0] - Q.A[0]) + pp.X[0] - qq.X[0];
  s.PQy = (P.A[1] - Q.A[1]) + pp.X[1] - qq.X[1];
  s.PQz = (P.A[2] - Q.A[2]) + pp.X[2] - qq.X[2];
  s.PAx = pp.X[0]; s.PAy = pp.X[1]; s.PAz = pp.X[2];
  s.QCx = qq.X[0]; s.QCy = qq.X[1]; s.QCz = qq.X[2];
  s.T = rho * (s.PQx * s.PQx + s.PQy * s.PQy + s.PQz * s.PQz);
  s.pref = pp.Kg * qq.Kg * sq;
  s.rz = eta * ize;
  s.re = zeta * ize;
  s.h2z = pp.h;
  s.h2e = qq.h;
  s.h2ze = 0.5 * ize;
  return s;
}
// gradient records (PrimPair): absolute product centres
__device__ __forceinline__ HLPrim hl_prim(const ShellPair& P, const ShellPair& Q, const PrimPair& pp,
                                          const PrimPair& qq) {
  HLPrim s;
  const double zeta = pp.g, eta = qq.g, ze = zeta + eta;
  const double sq = rsq64(ze), ize = sq * sq, rho = zeta * eta * ize;
  s.PQx = pp.P[0] - qq.P[0]; s.PQy = pp.P[1] - qq.P[1]; s.PQz = pp.P[2] - qq.P[2];
  s.PAx = pp.P[0] - P.A[0]; s.PAy = pp.P[1] - P.A[1]; s.PAz = pp.P[2] - P.A[2];
  s.QCx = qq.P[0] - Q.A[0]; s.QCy = qq.P[1] - Q.A[1]; s.QCz = qq.P[2] - Q.A[2];
  s.T = rho * (s.PQx * s.PQx + s.PQy * s.PQy + s.PQz * s.PQz);
  s.pref = pp.Kg * qq.Kg * sq;
  s.rz = eta * ize;
  s.re = zeta * ize;
  s.h2z = pp.h;
  s.h2e = qq.h;
  s.h2ze = 0.5 * ize;
  return s;
}

// 2D integrals of one primitive quartet by the vertical recursion: lane
// 3 r + d < 3 nr fills G[(r*3 + d)][i][k], i < ni <= 8, k < nk, for root r
// and axis d (the x table carries the prefactor and the weight); the last
// two ket columns stay in registers.
__device__ __forceinline__ void hl_vrr_axis(const HLDims& h, const HLPrim& s, int d, double t2, double wr, double* g) {
  const double B00 = s.h2ze * t2, B10 = s.h2z - s.h2z * s.rz * t2, B01 = s.h2e - s.h2e * s.re * t2;
  const double pa = d == 0 ? s.PAx : (d == 1 ? s.PAy : s.PAz);
  const double pq = d == 0 ? s.PQx : (d == 1 ? s.PQy : s.PQz);
  const double qc = d == 0 ? s.QCx : (d == 1 ? s.QCy : s.QCz);
  const double C00 = pa - s.rz * t2 * pq;
  const double D00 = qc + s.re * t2 * pq;
  const int ni = h.ni, nk = h.nk;
  double c1[8], c2[8];
  c1[0] = (d == 0) ? s.pref * wr : 1.0;
  g[0] = c1[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    double v = C00 * c1[i - 1];
    if (i > 1) v += (double)(i - 1) * B10 * c1[i - 2];
    c1[i] = v;
    if (i < ni) g[i * nk] = v;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) c2[i] = 0.0;
  for (int k = 1; k < nk; ++k) {
    const double bk = (double)(k - 1) * B01;
    double cn[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double v = D00 * c1[i] + bk * c2[i];
      if (i > 0) v += (double)i * B00 * c1[i - 1];
      cn[i] = v;
      if (i < ni) g[i * nk + k] = v;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      c2[i] = c1[i];
      c1[i] = cn[i];
    }
  }
}
// AXES: lane 3 r + d takes root r, axis d (shorter chain, more registers);
// otherwise lane r takes the three axes of root r one after another
template <bool AXES = true>
__device__ __forceinline__ void hl_vrr(const HLDims& h, const HLPrim& s, const double* __restrict__ rys, double* G) {
  const int lane = threadIdx.x;
  if constexpr (!AXES) {
    if (lane >= h.nr) return;
    double t2, wr;
    hl_root(h.nr, s.T, rys, lane, t2, wr);
    const int nrk = h.ni * h.nk;
    double* g = G + 3 * lane * nrk;
    hl_vrr_axis(h, s, 0, t2, wr, g);
    hl_vrr_axis(h, s, 1, t2, wr, g + nrk);
    hl_vrr_axis(h, s, 2, t2, wr, g + 2 * nrk);
    return;
  }
  if (lane >= 3 * h.nr) return;
  const int r = lane / 3, d = lane - 3 * r;
  double t2, wr;
  hl_root(h.nr, s.T, rys, r, t2, wr);
  const int nrk = h.ni * h.nk;
  hl_vrr_axis(h, s, d, t2, wr, G + (3 * r + d) * nrk);
}

// one index of a 4-index block through the canonical-Cartesian -> AO
// transform of level L: dst[l][o][r] = sum_i X[o][i] src[l][i][r]
__device__ __forceinline__ void hl_xform(const double* src, double* dst, int left, int ni, int no, int right, int L,
                                         int pure, const double* __restrict__ xf) {
  const double* X = xf + ((L - 2) * 2 + pure) * 100;
  const int n = left * no * right;
  for (int idx = threadIdx.x; idx < n; idx += HL_BLOCK) {
    const int l = idx / (no * right), rem = idx - l * no * right, o = rem / right, r = rem - o * right;
    double v = 0.0;
    for (int i = 0; i < ni; ++i) v += X[o * 10 + i] * src[(l * ni + i) * right + r];
    dst[idx] = v;
  }
}

// the same for a transform that keeps the component count (Cartesian
// shells: order and norms), in place: one lane per line (l, r), which it
// reads completely before writing it back
__device__ __forceinline__ void hl_xform_inplace(double* buf, int left, int n, int right, int L, int pure,
                                                 const double* __restrict__ xf) {
  const double* X = xf + ((L - 2) * 2 + pure) * 100;
  const int nl = left * right;
  for (int line = threadIdx.x; line < nl; line += HL_BLOCK) {
    const int l = line / right, r = line - l * right;
    double t[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) t[i] = i < n ? buf[(l * n + i) * right + r] : 0.0;
#pragma unroll
    for (int o = 0; o < 10; ++o) {
      if (o >= n) break;
      double v = 0.0;
#pragma unroll
      for (int i = 0; i < 10; ++i) v += X[o * 10 + i] * t[i];
      buf[(l * n + o) * right + r] = v;
    }
  }
}

// Final AO block [oa][ob][oc][od] of the quartet (P|Q) in LDS (the returned
// pointer); nprim = surviving primitive quartets, 0 when none survived (the
// block is then not formed).  Primitive screens as contract(): |K| < qcut
// per pair and K_ab^2 K_cd^2 < qcut^2 (zeta + eta) per quartet.
template <bool AXES = true>
__device__ double* hl_block(const HLDims& h, const HLTabs& tb, const ShellPair& P, const ShellPair& Q,
                            const PrimJK* prims, const double* __restrict__ rys, const double* __restrict__ xf,
                            double qcut, double imu2, double* lds, int& nprim) {
  const int lane = threadIdx.x;
  double* G = lds + h.offG;
  double* bufA = lds + h.offA;
  double* bufB = lds + h.offB;
  double* C = lds + h.offC;
  const int nef = h.ne * h.nf, nrk = h.ni * h.nk, nk = h.nk, nf = h.nf;
  __syncthreads();  // the previous quartet's readers of LDS are done
  for (int i = lane; i < nef; i += HL_BLOCK) bufA[i] = 0.0;
  hl_coef(P, Q, C);
  nprim = 0;
  const double qcut2 = qcut * qcut;
  for (int iq = 0; iq < Q.nprim; ++iq) {
    const PrimJK qq = load_uniform(prims, Q.prim0 + iq);
    if (fabs(qq.Kg) < qcut) continue;
    for (int ip = 0; ip < P.nprim; ++ip) {
      const PrimJK pp = load_uniform(prims, P.prim0 + ip);
      if (fabs(pp.Kg) < qcut) continue;
      const double fqz = pp.Kg * qq.Kg;
      if (fqz * fqz < qcut2 * (pp.g + qq.g + pp.g * qq.g * imu2)) continue;
      ++nprim;
      const HLPrim s = hl_prim(P, Q, pp, qq, imu2);
      __syncthreads();
      hl_vrr<AXES>(h, s, rys, G);
      __syncthreads();
      HLStep st(lane, nf);
      for (int idx = lane; idx < nef; idx += HL_BLOCK, st.next()) {
        const int pe = tb.eT[st.i], pf = tb.fT[st.j];
        const int ox = (pe & 15) * nk + (pf & 15), oy = nrk + ((pe >> 4) & 15) * nk + ((pf >> 4) & 15),
                  oz = 2 * nrk + (pe >> 8) * nk + (pf >> 8);
        double v = 0.0;
        for (int r = 0; r < h.nr; ++r) {
          const double* g = G + 3 * r * nrk;
          v += g[ox] * g[oy] * g[oz];
        }
        bufA[idx] += v;
      }
    }
  }
  if (nprim == 0) return nullptr;
  __syncthreads();
  const double* cab = C;
  const double* ccd = C + 75;
  // bra HRR: Y[a][b][f] = sum_j prod_d C(b_d, j_d) AB_d^(b_d - j_d) acc[a + j][f]
  {
    const int n = h.na * h.nb * nf;
    HLStep st(lane, nf);
    for (int idx = lane; idx < n; idx += HL_BLOCK, st.next()) {
      const int p = tb.abT[st.i], f = st.j;
      const int ax = (p >> 8) & 15, ay = (p >> 12) & 15, az = (p >> 16) & 15;
      const int bx = (p >> 20) & 15, by = (p >> 24) & 15, bz = (p >> 28) & 15;
      double v = 0.0;
      for (int jx = 0; jx <= bx; ++jx)
        for (int jy = 0; jy <= by; ++jy) {
          const double cxy = cab[bx * 5 + jx] * cab[25 + by * 5 + jy];
          for (int jz = 0; jz <= bz; ++jz)
            v += cxy * cab[50 + bz * 5 + jz] * bufA[hl_eidx(h.la, ax + jx, ay + jy, az + jz) * nf + f];
        }
      bufB[idx] = v;
    }
  }
  double* cur = bufB;
  double* oth = bufA;
  auto flip = [&]() {
    double* t = cur;
    cur = oth;
    oth = t;
  };
  int ob = h.nb, oa = h.na;
  if (h.lb >= 2) {
    __syncthreads();
    if (h.ob == h.nb) {
      hl_xform_inplace(cur, h.na, h.nb, nf, h.lb, h.pure, xf);
    } else {
      hl_xform(cur, oth, h.na, h.nb, h.ob, nf, h.lb, h.pure, xf);
      flip();
    }
    ob = h.ob;
  }
  if (h.la >= 2) {
    __syncthreads();
    if (h.oa == h.na) {
      hl_xform_inplace(cur, 1, h.na, ob * nf, h.la, h.pure, xf);
    } else {
      hl_xform(cur, oth, 1, h.na, h.oa, ob * nf, h.la, h.pure, xf);
      flip();
    }
    oa = h.oa;
  }
  // ket HRR: W[ab][c][d] = sum_m prod_d C(d_d, m_d) CD_d^(d_d - m_d) Y[ab][c + m]
  __syncthreads();
  {
    const int ncd = h.nc * h.nd, n = oa * ob * ncd;
    HLStep st(lane, ncd);
    for (int idx = lane; idx < n; idx += HL_BLOCK, st.next()) {
      const int p = tb.cdT[st.j];
      const int cx = (p >> 8) & 15, cy = (p >> 12) & 15, cz = (p >> 16) & 15;
      const int dx = (p >> 20) & 15, dy = (p >> 24) & 15, dz = (p >> 28) & 15;
      const double* y = cur + st.i * nf;
      double v = 0.0;
      for (int mx = 0; mx <= dx; ++mx)
        for (int my = 0; my <= dy; ++my) {
          const double cxy = ccd[dx * 4 + mx] * ccd[16 + dy * 4 + my];
          for (int mz = 0; mz <= dz; ++mz)
            v += cxy * ccd[32 + dz * 4 + mz] * y[hl_eidx(h.lc, cx + mx, cy + my, cz + mz)];
        }
      oth[idx] = v;
    }
    flip();
  }
  if (h.ld >= 2) {
    __syncthreads();
    if (h.od == h.nd) {
      hl_xform_inplace(cur, oa * ob * h.nc, h.nd, 1, h.ld, h.pure, xf);
    } else {
      hl_xform(cur, oth, oa * ob * h.nc, h.nd, h.od, 1, h.ld, h.pure, xf);
      flip();
    }
  }
  if (h.lc >= 2) {
    __syncthreads();
    if (h.oc == h.nc) {
      hl_xform_inplace(cur, oa * ob, h.nc, h.od, h.lc, h.pure, xf);
    } else {
      hl_xform(cur, oth, oa * ob, h.nc, h.oc, h.od, h.lc, h.pure, xf);
      flip();
    }
  }
  __syncthreads();
  return cur;
}

// ------------------------------------------------------------ Fock build
// Wavefront tasks as k_jk: a block takes one 64-ket chunk of one segment,
// screens its kets lane-parallel with OpenQP's tests, then evaluates and
// digests the surviving quartets one after another, all 64 lanes on each.
// Response-pass digestion of one quartet (ab|cd) for every non-symmetric
// density M_v (int2_mrsf_data_t_update, tdhf_mrsf_lib.F90:278-308): Coulomb
// with the symmetrised blocks M + M^T into F_ab, F_ba, F_cd, F_dc; exchange
// into the eight targets F_ac, F_ca, F_ad, F_da, F_bc, F_cb, F_bd, F_db, each
// from its own block of M (F_ac from M_bd, F_ca from M_db, ...).  deg holds
// the shell-level degeneracy factors; the outputs are full nbf x nbf.  The
// lanes run over (density, output element) pairs and read M from memory, so
// small classes keep the wavefront busy across the densities.
__device__ __forceinline__ void hl_general(const HLDims& h, const double* V, const ShellPair& P, const ShellPair& Q,
                                           double deg, const JKParams& prm, int lane) {
  const int nbf = prm.nbf, nd = prm.gnd;
  const size_t n2 = (size_t)nbf * nbf;
  const int oa = P.ao_a, ob = P.ao_b, oc = Q.ao_a, od = Q.ao_b;
  const int OA = h.oa, OB = h.ob, OC = h.oc, OD = h.od;
  const double cut = prm.cutoff;
  const double cj = deg * prm.sc, ck = -deg * prm.se;
  auto val = [&](int a, int b, int c, int d) {
    const double v = V[((a * OB + b) * OC + c) * OD + d];
    return fabs(v) < cut ? 0.0 : v;  // OpenQP element cutoff on the AO value
  };
  auto put = [&](double* Fv, int x, int y, double w) {
    if (w != 0.0) fatomic(Fv + (size_t)x * nbf + y, w);
  };
  // Coulomb
  for (int t = lane; t < nd * OA * OB; t += HL_BLOCK) {
    const int v = t / (OA * OB), e = t - v * (OA * OB);
    if (!prm.gcoul[v]) continue;
    const int a = e / OB, b = e - a * OB;
    const double* M = prm.Dg + (size_t)v * n2;
    double s = 0.0;
    for (int c = 0; c < OC; ++c)
      for (int d = 0; d < OD; ++d)
        s += val(a, b, c, d) * (M[(size_t)(oc + c) * nbf + od + d] + M[(size_t)(od + d) * nbf + oc + c]);
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, oa + a, ob + b, cj * s);
    put(Fv, ob + b, oa + a, cj * s);
  }
  for (int t = lane; t < nd * OC * OD; t += HL_BLOCK) {
    const int v = t / (OC * OD), e = t - v * (OC * OD);
    if (!prm.gcoul[v]) continue;
    const int c = e / OD, d = e - c * OD;
    const double* M = prm.Dg + (size_t)v * n2;
    double s = 0.0;
    for (int a = 0; a < OA; ++a)
      for (int b = 0; b < OB; ++b)
        s += val(a, b, c, d) * (M[(size_t)(oa + a) * nbf + ob + b] + M[(size_t)(ob + b) * nbf + oa + a]);
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, oc + c, od + d, cj * s);
    put(Fv, od + d, oc + c, cj * s);
  }
  // exchange: F_ac / F_ca
  for (int t = lane; t < nd * OA * OC; t += HL_BLOCK) {
    const int v = t / (OA * OC), e = t - v * (OA * OC);
    const int a = e / OC, c = e - a * OC;
    const double* M = prm.Dg + (size_t)v * n2;
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < OB; ++b)
      for (int d = 0; d < OD; ++d) {
        const double w = val(a, b, c, d);
        s1 += w * M[(size_t)(ob + b) * nbf + od + d];
        s2 += w * M[(size_t)(od + d) * nbf + ob + b];
      }
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, oa + a, oc + c, ck * s1);
    put(Fv, oc + c, oa + a, ck * s2);
  }
  // F_ad / F_da
  for (int t = lane; t < nd * OA * OD; t += HL_BLOCK) {
    const int v = t / (OA * OD), e = t - v * (OA * OD);
    const int a = e / OD, d = e - a * OD;
    const double* M = prm.Dg + (size_t)v * n2;
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < OB; ++b)
      for (int c = 0; c < OC; ++c) {
        const double w = val(a, b, c, d);
        s1 += w * M[(size_t)(ob + b) * nbf + oc + c];
        s2 += w * M[(size_t)(oc + c) * nbf + ob + b];
      }
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, oa + a, od + d, ck * s1);
    put(Fv, od + d, oa + a, ck * s2);
  }
  // F_bc / F_cb
  for (int t = lane; t < nd * OB * OC; t += HL_BLOCK) {
    const int v = t / (OB * OC), e = t - v * (OB * OC);
    const int b = e / OC, c = e - b * OC;
    const double* M = prm.Dg + (size_t)v * n2;
    double s1 = 0.0, s2 = 0.0;
    for (int a = 0; a < OA; ++a)
      for (int d = 0; d < OD; ++d) {
        const double w = val(a, b, c, d);
        s1 += w * M[(size_t)(oa + a) * nbf + od + d];
        s2 += w * M[(size_t)(od + d) * nbf + oa + a];
      }
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, ob + b, oc + c, ck * s1);
    put(Fv, oc + c, ob + b, ck * s2);
  }
  // F_bd / F_db
  for (int t = lane; t < nd * OB * OD; t += HL_BLOCK) {
    const int v = t / (OB * OD), e = t - v * (OB * OD);
    const int b = e / OD, d = e - b * OD;
    const double* M = prm.Dg + (size_t)v * n2;
    double s1 = 0.0, s2 = 0.0;
    for (int a = 0; a < OA; ++a)
      for (int c = 0; c < OC; ++c) {
        const double w = val(a, b, c, d);
        s1 += w * M[(size_t)(oa + a) * nbf + oc + c];
        s2 += w * M[(size_t)(oc + c) * nbf + oa + a];
      }
    double* Fv = prm.Fg + (size_t)v * n2;
    put(Fv, ob + b, od + d, ck * s1);
    put(Fv, od + d, ob + b, ck * s2);
  }
}

// One 64-ket chunk (wave) of one class's wavefront tasks, by one block.
// GEN: the response pass over non-symmetric densities (prm.Dg), a separate
// instance so the Fock-build kernels keep their register allocation.
template <bool GEN>
__device__ __forceinline__ void jkh_chunk(const JKParams& prm, const WaveWork& ww, const HLDims& h, long long blk,
                                          double* lds) {
  const HLTabs tb = hl_tables(h, lds);
  // block blk takes kets [sub * kpb, (sub + 1) * kpb) of 64-ket chunk `wave`
  const int kpb = ww.kpb > 0 ? ww.kpb : 64, per = 64 / kpb;
  const long long wave = blk / per;
  const int sub = (int)(blk - wave * per);
  if (wave >= ww.nwaves) return;  // block-uniform
  int lo = 0, hi = ww.nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
  }
  const int seg = __builtin_amdgcn_readfirstlane(lo);
  const int lane = threadIdx.x;
  const int k = (int)(wave - ww.choff[seg]) * 64 + sub * kpb + lane;
  const bool active = lane < kpb && k < ww.seg_n[seg];
  const int ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
  const ShellPair P = prm.pairs[ip];
  int iq = 0;
  bool skip = true, mine = true;
  if (active) {
    iq = ww.ket[ww.seg_k0[seg] + k];
    const ShellPair Q = prm.pairs[iq];
    const bool p_first = (P.si > Q.si) || (P.si == Q.si && P.sj >= Q.sj);
    if (prm.own) mine = prm.own[p_first ? ip : iq] != 0;
    const double qbra = p_first ? P.Q : Q.Q;
    const int n = prm.nsh;
    const double* ds = prm.dsh;
    double dm = fmax(4.0 * ds[P.si * n + P.sj], 4.0 * ds[Q.si * n + Q.sj]);
    dm = fmax(dm, fmax(ds[P.sj * n + Q.sj], ds[P.sj * n + Q.si]));
    dm = fmax(dm, fmax(ds[P.si * n + Q.sj], ds[P.si * n + Q.si]));
    skip = !mine || (qbra * prm.maxden < prm.cutoff) || (P.Q * Q.Q * dm < prm.cutoff);
  }
  unsigned long long* ctr = prm.counters + 8 * (int)(wave & (OQPH_CTR_SLOTS - 1));
  count_wave(ctr + 1, active && mine && skip);
  count_wave(ctr + 0, !skip);
  unsigned long long todo = __ballot(!skip);
  const long long fprim = flops_prim_rt(h.la, h.lb, h.lc, h.ld);
  const long long nout = (long long)h.oa * h.ob * h.oc * h.od;
  unsigned long long flops = 0ull, prims_used = 0ull;
  const int oa = P.ao_a, ob = P.ao_b;
  const double* Dp = prm.D;
  double* F = prm.F;
  const double cut = prm.cutoff;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    todo &= todo - 1ull;
    const int jq = __builtin_amdgcn_readfirstlane(__shfl(iq, l));
    const ShellPair Q = prm.pairs[jq];
    int nprim = 0;
    // the response pass keeps one lane per root (its densities' digestion
    // needs the registers)
    const double* V = hl_block<!GEN>(h, tb, P, Q, prm.prims, prm.rys, prm.xform, 1.0e-4 * cut, prm.inv_mu2, lds, nprim);
    prims_used += (unsigned long long)nprim;
    if (nprim == 0) continue;
    flops += (unsigned long long)(nprim * fprim + (GEN ? 20 * prm.gnd : prm.F2 ? 24 : 12) * nout);
    double deg = 1.0;
    if (P.sa == P.sb) deg *= 0.5;
    if (Q.sa == Q.sb) deg *= 0.5;
    if (ip == jq) deg *= 0.5;
    if constexpr (GEN) {
      hl_general(h, V, P, Q, deg, prm, lane);
      continue;
    }
    const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se;
    const int oc = Q.ao_a, od = Q.ao_b;
    const int OA = h.oa, OB = h.ob, OC = h.oc, OD = h.od;
    // the six density blocks of this quartet, staged in LDS (block stride
    // h.dstr = the square of the class's largest AO count)
    double* sD = lds + h.offD;
    const int ds = h.dstr;
    for (int t = lane; t < 6 * ds; t += HL_BLOCK) {
      const int blk = t / ds, e = t - ds * blk;
      const int r0 = blk == 0 ? oa : blk == 1 ? oc : blk <= 3 ? oa : ob;
      const int c0 = blk == 0 ? ob : blk == 1 ? od : (blk == 2 || blk == 4) ? oc : od;
      const int nc = blk == 0 ? OB : blk == 1 ? OD : (blk == 2 || blk == 4) ? OC : OD;
      const int nr = blk == 0 ? OA : blk == 1 ? OC : blk <= 3 ? OA : OB;
      if (e < nr * nc) sD[t] = Dp[tri_idx(r0 + e / nc, c0 + e % nc)];
    }
    __syncthreads();
    const double *Dab = sD, *Dcd = sD + ds, *Dac = sD + 2 * ds, *Dad = sD + 3 * ds, *Dbc = sD + 4 * ds,
                 *Dbd = sD + 5 * ds;
    auto val = [&](int a, int b, int c, int d) {
      const double v = V[((a * OB + b) * OC + c) * OD + d];
      return fabs(v) < cut ? 0.0 : v;  // OpenQP element cutoff on the AO value
    };
    auto put = [&](double* Ft, int x, int y, double v) {
      v *= (x == y ? 2.0 : 1.0);
      if (v != 0.0) fatomic(Ft + tri_idx(x, y), v);
    };
    // J_ab and J_cd
    for (int t = lane; t < OA * OB; t += HL_BLOCK) {
      const int a = t / OB, b = t - a * OB;
      double s = 0.0;
      for (int c = 0; c < OC; ++c)
        for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dcd[c * OD + d];
      put(F, oa + a, ob + b, cj * s);
    }
    for (int t = lane; t < OC * OD; t += HL_BLOCK) {
      const int c = t / OD, d = t - c * OD;
      double s = 0.0;
      for (int a = 0; a < OA; ++a)
        for (int b = 0; b < OB; ++b) s += val(a, b, c, d) * Dab[a * OB + b];
      put(F, oc + c, od + d, cj * s);
    }
    // K_ac += V D_bd, K_ad += V D_bc, K_bc += V D_ad, K_bd += V D_ac
    auto kblocks = [&](double* Ft) {
      for (int t = lane; t < OA * OC; t += HL_BLOCK) {
        const int a = t / OC, c = t - a * OC;
        double s = 0.0;
        for (int b = 0; b < OB; ++b)
          for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dbd[b * OD + d];
        put(Ft, oa + a, oc + c, ck * s);
      }
      for (int t = lane; t < OA * OD; t += HL_BLOCK) {
        const int a = t / OD, d = t - a * OD;
        double s = 0.0;
        for (int b = 0; b < OB; ++b)
          for (int c = 0; c < OC; ++c) s += val(a, b, c, d) * Dbc[b * OC + c];
        put(Ft, oa + a, od + d, ck * s);
      }
      for (int t = lane; t < OB * OC; t += HL_BLOCK) {
        const int b = t / OC, c = t - b * OC;
        double s = 0.0;
        for (int a = 0; a < OA; ++a)
          for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dad[a * OD + d];
        put(Ft, ob + b, oc + c, ck * s);
      }
      for (int t = lane; t < OB * OD; t += HL_BLOCK) {
        const int b = t / OD, d = t - b * OD;
        double s = 0.0;
        for (int a = 0; a < OA; ++a)
          for (int c = 0; c < OC; ++c) s += val(a, b, c, d) * Dac[a * OC + c];
        put(Ft, ob + b, od + d, ck * s);
      }
    };
    kblocks(F);
    if (prm.F2) {  // launch-uniform: the open-shell pass, D2's exchange blocks into F2
      __syncthreads();
      for (int t = lane + 2 * ds; t < 6 * ds; t += HL_BLOCK) {
        const int blk = t / ds, e = t - ds * blk;
        const int r0 = blk <= 3 ? oa : ob;
        const int c0 = (blk == 2 || blk == 4) ? oc : od;
        const int nc = (blk == 2 || blk == 4) ? OC : OD;
        const int nr = blk <= 3 ? OA : OB;
        if (e < nr * nc) sD[t] = prm.D2[tri_idx(r0 + e / nc, c0 + e % nc)];
      }
      __syncthreads();
      kblocks(prm.F2);
    }
  }
  if (lane == 0) {
    if (flops) atomicAdd(ctr + 3, flops);
    if (prims_used) atomicAdd(ctr + 4, prims_used);
  }
}

template <bool GEN>
__global__ __launch_bounds__(HL_BLOCK) __attribute__((amdgpu_waves_per_eu(OQPH_HL_WAVES))) void k_jkh(JKParams prm, WaveWork ww, HLDims h) {
  extern __shared__ double lds[];
  jkh_chunk<GEN>(prm, ww, h, (long long)blockIdx.x, lds);
}

// Several run-time classes in one launch: block b takes wave b - wave0 of the
// task (class) whose range holds it; the tasks come sorted by decreasing
// cost per wave, so the most expensive chunks start first and the launch ends
// on the cheapest.  The task tables hold offsets; the list bases are launch
// arguments (the lists may live in different buffers from build to build).
template <bool GEN>
__global__ __launch_bounds__(HL_BLOCK) __attribute__((amdgpu_waves_per_eu(OQPH_HL_WAVES))) void k_jkh_multi(
    JKParams prm, HLTaskList tl) {
  extern __shared__ double lds[];
  const long long b = blockIdx.x;
  int lo = 0, hi = tl.ntask;
  while (hi - lo > 1) {  // scalar binary search over the task prefix
    const int mid = (lo + hi) >> 1;
    if (tl.task[mid].wave0 <= b) lo = mid; else hi = mid;
  }
  const int t = __builtin_amdgcn_readfirstlane(lo);
  const HLTask& tk = tl.task[t];
  WaveWork ww;
  ww.seg_bra = tl.seg + tk.seg0;
  ww.seg_k0 = tl.seg + tl.nseg + tk.seg0;
  ww.seg_n = tl.seg + 2 * tl.nseg + tk.seg0;
  ww.choff = tl.choff + tk.ch0;
  ww.ket = tl.ket + tk.ket0;
  ww.nseg = tk.nseg;
  ww.nwaves = tk.nwaves;
  ww.kpb = tk.kpb;
  const HLDims h = tk.h;
  jkh_chunk<GEN>(prm, ww, h, b - tk.wave0, lds);
}

// --------------------------------------------------------- Schwarz bounds
// One block per pair of the list: Q_ab = sqrt(max |(ab|ab)|) over the AO block.
__global__ __launch_bounds__(HL_BLOCK) void k_schwarzh(const ShellPair* pairs, const PrimJK* prims, const int* list,
                                                       int n, const double* rys, const double* xf, double qcut,
                                                       double imu2, double* Qout, HLDims h) {
  extern __shared__ double lds[];
  const HLTabs tb = hl_tables(h, lds);
  if ((int)blockIdx.x >= n) return;
  const int ip = list[blockIdx.x];
  const ShellPair P = pairs[ip];
  int nprim = 0;
  const double* V = hl_block(h, tb, P, P, prims, rys, xf, qcut, imu2, lds, nprim);
  double m = 0.0;
  if (nprim) {
    const int nv = h.oa * h.ob * h.oc * h.od;
    for (int i = threadIdx.x; i < nv; i += HL_BLOCK) m = fmax(m, fabs(V[i]));
  }
  m = wave_max(m);
  if (threadIdx.x == 0) Qout[ip] = sqrt(m);
}

// --------------------------------------------------------------- gradient
// Gradient wavefront tasks (coarse-screened lists): a block takes one chunk
// of kets of one bra; per quartet the lanes form the Cartesian Gamma block
// in LDS and apply the fine screen; per primitive quartet and root they
// transfer the 2D integrals to every (a, b, c, d) exponent the derivatives
// need (T, LDS) and contract them with their Gamma entries; per-atom sums
// go through LDS.
// One chunk (wave) of one gradient class's wavefront tasks, by one block.
// MRSF: Gamma from the MRSF matrices (a separate instance, so the
// ground-state kernels keep their register allocation)
template <bool MRSF>
__device__ __forceinline__ void gradh_chunk(const GradParams& prm, const WaveWork& ww, const HLDims& h, long long blk,
                                            double* lds) {
  const HLTabs tb = hl_tables(h, lds);
  double* s_de = lds + h.offB;  // 3 natm
  for (int i = threadIdx.x; i < 3 * prm.natm; i += HL_BLOCK) s_de[i] = 0.0;
  __syncthreads();
  // block blk takes kets [sub * kpb, (sub + 1) * kpb) of 64-ket chunk `wave`
  const int kpb = ww.kpb > 0 ? ww.kpb : 64, per = 64 / kpb;
  const long long wave = blk / per;
  const int sub = (int)(blk - wave * per);
  if (wave >= ww.nwaves) return;  // block-uniform
  int lo = 0, hi = ww.nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
  }
  const int seg = __builtin_amdgcn_readfirstlane(lo);
  const int lane = threadIdx.x;
  const int k = (int)(wave - ww.choff[seg]) * 64 + sub * kpb + lane;
  const bool active = lane < kpb && k < ww.seg_n[seg];
  const int ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
  const ShellPair P = prm.pairs[ip];
  const int iq = active ? ww.ket[ww.seg_k0[seg] + k] : 0;
  unsigned long long* ctr = prm.counters + 8 * (int)(blk & (OQPH_CTR_SLOTS - 1));
  unsigned long long todo = __ballot(active);
  double* G = lds + h.offG;
  double* Gam = lds + h.offA;
  double* T = lds + h.offT;
  double* C = lds + h.offC;
  const double* cab = C;
  const double* ccd = C + 75;
  const int NCD = h.nc * h.nd, NG = h.na * h.nb * NCD;
  const int n = prm.ncf;
  const double* Dc = prm.Dc;
  const int oa = prm.coff[P.sa], ob = prm.coff[P.sb];
  const long long fprim = flops_grad_prim_rt(h.la, h.lb, h.lc, h.ld);
  unsigned long long ncomp = 0ull, nskip = 0ull, flops = 0ull, prims_used = 0ull;
  const int nrk = h.ni * h.nk, nk = h.nk, nt = h.nt, n4 = h.n4;
  double* R = lds + h.offR;
  const int NTb = h.lb + 2, NTc = h.lc + 2, NTd = h.ld + 1;
  const int sa = NTb * NTc * NTd, sb = NTc * NTd, sc = NTd;  // T strides of ia, ib, ic (id: 1)
  const int lab1 = h.la + h.lb + 1, lcd1 = h.lc + h.ld + 1;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    todo &= todo - 1ull;
    const int jq = __builtin_amdgcn_readfirstlane(__shfl(iq, l));
    const ShellPair Q = prm.pairs[jq];
    const int oc = prm.coff[Q.sa], od = prm.coff[Q.sb];
    __syncthreads();  // the previous quartet's readers of Gamma, C and T are done
    double gm = 0.0;
    {
      HLStep st(lane, NCD);
      for (int t = lane; t < NG; t += HL_BLOCK, st.next()) {
        const int pab = tb.abT[st.i], pcd = tb.cdT[st.j];
        const int a = pab & 15, b = (pab >> 4) & 15, c = pcd & 15, d = (pcd >> 4) & 15;
        double g;
        if constexpr (MRSF)
          g = mrsf_gamma(prm, oa + a, ob + b, oc + c, od + d);
        else
          g = 4.0 * prm.cs * Dc[(oa + a) * n + ob + b] * Dc[(oc + c) * n + od + d] -
              prm.hs * (Dc[(oa + a) * n + oc + c] * Dc[(ob + b) * n + od + d] +
                        Dc[(oa + a) * n + od + d] * Dc[(ob + b) * n + oc + c]);
        Gam[t] = g;
        gm = fmax(gm, fabs(g));
      }
    }
    gm = wave_max(gm);
    // fine screen on the weighted contribution (grd2.F90:348)
    if (gm * P.Q * Q.Q < 0.5 * prm.gcut) {
      ++nskip;
      continue;
    }
    ++ncomp;
    hl_coef(P, Q, C);
    const double dab2 = gm * gm;
    double f0x = 0, f0y = 0, f0z = 0, f1x = 0, f1y = 0, f1z = 0, f2x = 0, f2y = 0, f2z = 0;
    int nprim = 0;
    for (int iqq = 0; iqq < Q.nprim; ++iqq) {
      const PrimPair qq = load_uniform(prm.prims, Q.prim0 + iqq);
      for (int ipp = 0; ipp < P.nprim; ++ipp) {
        const PrimPair pp = load_uniform(prm.prims, P.prim0 + ipp);
        const double ze = pp.g + qq.g, fqz = pp.Kg * qq.Kg;
        // primitive screens (grd2_rys.F90:375-376), prefactor cutoff (1e-20)^2
        if (fqz * fqz < 1.0e-40 * ze) continue;
        if (fqz * fqz * dab2 < prm.dabcut2 * ze) continue;
        ++nprim;
        const HLPrim s = hl_prim(P, Q, pp, qq);
        __syncthreads();  // the previous primitive's readers of G and T are done
        hl_vrr(h, s, prm.rys, G);
        const double a2 = 2.0 * pp.a, b2 = 2.0 * pp.b, c2 = 2.0 * qq.a;
        for (int r = 0; r < h.nr; ++r) {
          __syncthreads();  // G written, or the previous root's readers of T done
          // T[x][ia][ib][ic][id] = sum_j sum_m C(ib,j) AB^(ib-j) C(id,m) CD^(id-m) G_x(ia+j, ic+m)
          {
            HLStep st(lane, nt);
            for (int idx = lane; idx < 3 * nt; idx += HL_BLOCK, st.next()) {
              const int x = st.i, pt = tb.tT[st.j];
              const int ia = pt & 15, ib = (pt >> 4) & 15, ic = (pt >> 8) & 15, id = (pt >> 12) & 15;
              double v = 0.0;
              if (ia + ib <= lab1 && ic + id <= lcd1) {
                const double* g = G + (3 * r + x) * nrk;
                const double* cb = cab + 25 * x + 5 * ib;
                const double* cd = ccd + 16 * x + 4 * id;
                for (int j = 0; j <= ib; ++j) {
                  double u = 0.0;
                  for (int m = 0; m <= id; ++m) u += cd[m] * g[(ia + j) * nk + ic + m];
                  v += cb[j] * u;
                }
              }
              T[idx] = v;
            }
          }
          __syncthreads();
          // per axis and exponent tuple (ia, ib, ic, id) of the components:
          // the transferred 2D integral and its A, B and C derivatives
          //   d/dA = 2 alpha (ia+1) - ia (ia-1)  (likewise B, C), stored as
          // four consecutive doubles, so a component reads two b128 words per
          // axis instead of gathering seven table entries
          for (int idx = lane; idx < 3 * n4; idx += HL_BLOCK) {
            const int x = idx / n4, t = idx - x * n4;
            const int id = t % (h.ld + 1), t1 = t / (h.ld + 1);
            const int ic = t1 % (h.lc + 1), t2 = t1 / (h.lc + 1);
            const int ib = t2 % (h.lb + 1), ia = t2 / (h.lb + 1);
            const double* tx = T + x * nt + ia * sa + ib * sb + ic * sc + id;
            double* rr = R + 4 * idx;
            rr[0] = tx[0];
            rr[1] = a2 * tx[sa] - (ia > 0 ? (double)ia * tx[-sa] : 0.0);
            rr[2] = b2 * tx[sb] - (ib > 0 ? (double)ib * tx[-sb] : 0.0);
            rr[3] = c2 * tx[sc] - (ic > 0 ? (double)ic * tx[-sc] : 0.0);
          }
          __syncthreads();
          HLStep st(lane, NCD);
          for (int t = lane; t < NG; t += HL_BLOCK, st.next()) {
            const double g = Gam[t];
            if (g == 0.0) continue;
            const int pab = tb.abT[st.i], pcd = tb.cdT[st.j];
            double I0[3], dA[3], dB[3], dC[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              const int ia = (pab >> (8 + 4 * x)) & 15, ib = (pab >> (20 + 4 * x)) & 15;
              const int ic = (pcd >> (8 + 4 * x)) & 15, id = (pcd >> (20 + 4 * x)) & 15;
              const int tt = ((ia * (h.lb + 1) + ib) * (h.lc + 1) + ic) * (h.ld + 1) + id;
              const double2 v0 = *(const double2*)(R + 4 * (x * n4 + tt));
              const double2 v1 = *(const double2*)(R + 4 * (x * n4 + tt) + 2);
              I0[x] = v0.x;
              dA[x] = v0.y;
              dB[x] = v1.x;
              dC[x] = v1.y;
            }
            const double gyz = g * I0[1] * I0[2], gxz = g * I0[0] * I0[2], gxy = g * I0[0] * I0[1];
            f0x += dA[0] * gyz; f0y += dA[1] * gxz; f0z += dA[2] * gxy;
            f1x += dB[0] * gyz; f1y += dB[1] * gxz; f1z += dB[2] * gxy;
            f2x += dC[0] * gyz; f2y += dC[1] * gxz; f2z += dC[2] * gxy;
          }
        }
      }
    }
    prims_used += (unsigned long long)nprim;
    flops += (unsigned long long)(nprim * fprim);
    if (nprim == 0) continue;
    double deg = 1.0;
    if (P.sa == P.sb) deg *= 0.5;
    if (Q.sa == Q.sb) deg *= 0.5;
    if (ip == jq) deg *= 0.5;
    const int at0 = prm.shell_atom[P.sa], at1 = prm.shell_atom[P.sb], at2 = prm.shell_atom[Q.sa],
              at3 = prm.shell_atom[Q.sb];
    const double fd[3][3] = {{f0x, f0y, f0z}, {f1x, f1y, f1z}, {f2x, f2y, f2z}};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double fa = wave_sum(fd[0][d]), fb = wave_sum(fd[1][d]), fc = wave_sum(fd[2][d]);
      if (lane == 0) {
        atomicAdd(&s_de[3 * at0 + d], deg * fa);
        atomicAdd(&s_de[3 * at1 + d], deg * fb);
        atomicAdd(&s_de[3 * at2 + d], deg * fc);
        atomicAdd(&s_de[3 * at3 + d], -deg * (fa + fb + fc));
      }
    }
  }
  if (lane == 0) {
    if (ncomp) atomicAdd(ctr + 0, ncomp);
    if (nskip) atomicAdd(ctr + 1, nskip);
    if (flops) atomicAdd(ctr + 3, flops);
    if (prims_used) atomicAdd(ctr + 4, prims_used);
  }
  __syncthreads();
  for (int i = lane; i < 3 * prm.natm; i += HL_BLOCK)
    if (s_de[i] != 0.0) unsafeAtomicAdd(prm.de + (size_t)(blockIdx.x & (OQPH_DE_SLOTS - 1)) * 3 * prm.natm + i, s_de[i]);
}

template <bool MRSF>
__global__ __launch_bounds__(HL_BLOCK) __attribute__((amdgpu_waves_per_eu(OQPH_HL_WAVES))) void k_gradh(GradParams prm, WaveWork ww, HLDims h) {
  extern __shared__ double lds[];
  gradh_chunk<MRSF>(prm, ww, h, (long long)blockIdx.x, lds);
}

// Several gradient run-time classes in one launch (as k_jkh_multi).
template <bool MRSF>
__global__ __launch_bounds__(HL_BLOCK) __attribute__((amdgpu_waves_per_eu(OQPH_HL_WAVES))) void k_gradh_multi(
    GradParams prm, HLTaskList tl) {
  extern __shared__ double lds[];
  const long long b = blockIdx.x;
  int lo = 0, hi = tl.ntask;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tl.task[mid].wave0 <= b) lo = mid; else hi = mid;
  }
  const int t = __builtin_amdgcn_readfirstlane(lo);
  const HLTask& tk = tl.task[t];
  WaveWork ww;
  ww.seg_bra = tl.seg + tk.seg0;
  ww.seg_k0 = tl.seg + tl.nseg + tk.seg0;
  ww.seg_n = tl.seg + 2 * tl.nseg + tk.seg0;
  ww.choff = tl.choff + tk.ch0;
  ww.ket = tl.ket + tk.ket0;
  ww.nseg = tk.nseg;
  ww.nwaves = tk.nwaves;
  ww.kpb = tk.kpb;
  const HLDims h = tk.h;
  gradh_chunk<MRSF>(prm, ww, h, b - tk.wave0, lds);
}

}  // namespace oqph
