// hl_device.h -- shell quartets of classes that contain an f shell (any
// L <= 3), evaluated by one wavefront per quartet with every intermediate in
// LDS, sized per class at launch (HLDims).  Covers the Fock build, the
// Schwarz bounds and the 2e gradient for those classes; the algorithm is the
// one of the register kernels (Rys quadrature, contracted (e0|f0), binomial
// HRR, AO transform, OpenQP element cutoff and digestion), with loop bounds
// read at run time instead of unrolled, so (ff|ff) needs no register file
// larger than the hardware's.
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

// ---------------------------------------------------------- small helpers
// Cartesian exponents of entry e of the level range that starts at L0
__device__ __forceinline__ void hl_comp(int L0, int e, int& x, int& y, int& z) {
  int L = L0;
  while (e >= ncart(L)) { e -= ncart(L); ++L; }
  x = c_lx(L, e);
  y = c_ly(L, e);
  z = c_lz(L, e);
}
// index of (x, y, z) inside the level range that starts at L0
__device__ __forceinline__ int hl_eidx(int L0, int x, int y, int z) {
  return ncum(x + y + z - 1) - ncum(L0 - 1) + c_idx(x, y, z);
}
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

// Scalars of one primitive quartet shared by every lane.
struct HLPrim {
  double PQ[3], PA[3], QC[3];
  double T, pref, rz, re, h2z, h2e, h2ze;
};
__device__ __forceinline__ HLPrim hl_prim(const ShellPair& P, const ShellPair& Q, const PrimPair& pp,
                                          const PrimPair& qq) {
  HLPrim s;
  const double zeta = pp.g, eta = qq.g, ze = zeta + eta;
  const double sq = rsq64(ze), ize = sq * sq, rho = zeta * eta * ize;
  double pq2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s.PQ[d] = pp.P[d] - qq.P[d];
    s.PA[d] = pp.P[d] - P.A[d];
    s.QC[d] = qq.P[d] - Q.A[d];
    pq2 += s.PQ[d] * s.PQ[d];
  }
  s.T = rho * pq2;
  s.pref = pp.Kg * qq.Kg * sq;
  s.rz = eta * ize;
  s.re = zeta * ize;
  s.h2z = pp.h;
  s.h2e = qq.h;
  s.h2ze = 0.5 * ize;
  return s;
}

// 2D integrals of one primitive quartet by the vertical recursion: lane
// r < nr fills G[(r*3 + d)][i][k], i < ni, k < nk, for root r and the three
// axes d (the x table carries the prefactor and the weight).
__device__ __forceinline__ void hl_vrr(const HLDims& h, const HLPrim& s, const double* __restrict__ rys, double* G) {
  const int r = threadIdx.x;
  if (r >= h.nr) return;
  double t2, wr;
  hl_root(h.nr, s.T, rys, r, t2, wr);
  const double B00 = s.h2ze * t2, B10 = s.h2z - s.h2z * s.rz * t2, B01 = s.h2e - s.h2e * s.re * t2;
  const int ni = h.ni, nk = h.nk;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double C00 = s.PA[d] - s.rz * t2 * s.PQ[d];
    const double D00 = s.QC[d] + s.re * t2 * s.PQ[d];
    double* g = G + (3 * r + d) * ni * nk;
    g[0] = (d == 0) ? s.pref * wr : 1.0;
    double gm2 = 0.0, gm1 = g[0];
    for (int i = 1; i < ni; ++i) {
      const double v = C00 * gm1 + (double)(i - 1) * B10 * gm2;
      g[i * nk] = v;
      gm2 = gm1;
      gm1 = v;
    }
    for (int k = 1; k < nk; ++k)
      for (int i = 0; i < ni; ++i) {
        double v = D00 * g[i * nk + k - 1];
        if (k > 1) v += (double)(k - 1) * B01 * g[i * nk + k - 2];
        if (i > 0) v += (double)i * B00 * g[(i - 1) * nk + k - 1];
        g[i * nk + k] = v;
      }
  }
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
__device__ double* hl_block(const HLDims& h, const ShellPair& P, const ShellPair& Q, const PrimPair* prims,
                            const double* __restrict__ rys, const double* __restrict__ xf, double qcut, double* lds,
                            int& nprim) {
  const int lane = threadIdx.x;
  double* G = lds + h.offG;
  double* bufA = lds + h.offA;
  double* bufB = lds + h.offB;
  const int nef = h.ne * h.nf, nrk = h.ni * h.nk;
  __syncthreads();  // the previous quartet's readers of LDS are done
  for (int i = lane; i < nef; i += HL_BLOCK) bufA[i] = 0.0;
  nprim = 0;
  const double qcut2 = qcut * qcut;
  for (int iq = 0; iq < Q.nprim; ++iq) {
    const PrimPair qq = load_uniform(prims, Q.prim0 + iq);
    if (fabs(qq.Kg) < qcut) continue;
    for (int ip = 0; ip < P.nprim; ++ip) {
      const PrimPair pp = load_uniform(prims, P.prim0 + ip);
      if (fabs(pp.Kg) < qcut) continue;
      const double fqz = pp.Kg * qq.Kg;
      if (fqz * fqz < qcut2 * (pp.g + qq.g)) continue;
      ++nprim;
      const HLPrim s = hl_prim(P, Q, pp, qq);
      __syncthreads();
      hl_vrr(h, s, rys, G);
      __syncthreads();
      for (int idx = lane; idx < nef; idx += HL_BLOCK) {
        const int e = idx / h.nf, f = idx - e * h.nf;
        int ex, ey, ez, fx, fy, fz;
        hl_comp(h.la, e, ex, ey, ez);
        hl_comp(h.lc, f, fx, fy, fz);
        double v = 0.0;
        for (int r = 0; r < h.nr; ++r) {
          const double* g = G + 3 * r * nrk;
          v += g[ex * h.nk + fx] * g[nrk + ey * h.nk + fy] * g[2 * nrk + ez * h.nk + fz];
        }
        bufA[idx] += v;
      }
    }
  }
  if (nprim == 0) return nullptr;
  __syncthreads();
  double AB[3], CD[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = P.A[d] - P.B[d];
    CD[d] = Q.A[d] - Q.B[d];
  }
  // bra HRR: Y[a][b][f] = sum_j prod_d C(b_d, j_d) AB_d^(b_d - j_d) acc[a + j][f]
  {
    const int n = h.na * h.nb * h.nf;
    for (int idx = lane; idx < n; idx += HL_BLOCK) {
      const int a = idx / (h.nb * h.nf), rem = idx - a * h.nb * h.nf, b = rem / h.nf, f = rem - b * h.nf;
      const int ax = c_lx(h.la, a), ay = c_ly(h.la, a), az = c_lz(h.la, a);
      const int bx = c_lx(h.lb, b), by = c_ly(h.lb, b), bz = c_lz(h.lb, b);
      double v = 0.0;
      for (int jx = 0; jx <= bx; ++jx)
        for (int jy = 0; jy <= by; ++jy)
          for (int jz = 0; jz <= bz; ++jz) {
            const double cf = (double)(binom(bx, jx) * binom(by, jy) * binom(bz, jz)) * hl_pow(AB[0], bx - jx) *
                              hl_pow(AB[1], by - jy) * hl_pow(AB[2], bz - jz);
            v += cf * bufA[hl_eidx(h.la, ax + jx, ay + jy, az + jz) * h.nf + f];
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
      hl_xform_inplace(cur, h.na, h.nb, h.nf, h.lb, h.pure, xf);
    } else {
      hl_xform(cur, oth, h.na, h.nb, h.ob, h.nf, h.lb, h.pure, xf);
      flip();
    }
    ob = h.ob;
  }
  if (h.la >= 2) {
    __syncthreads();
    if (h.oa == h.na) {
      hl_xform_inplace(cur, 1, h.na, ob * h.nf, h.la, h.pure, xf);
    } else {
      hl_xform(cur, oth, 1, h.na, h.oa, ob * h.nf, h.la, h.pure, xf);
      flip();
    }
    oa = h.oa;
  }
  // ket HRR: W[ab][c][d] = sum_m prod_d C(d_d, m_d) CD_d^(d_d - m_d) Y[ab][c + m]
  __syncthreads();
  {
    const int ncd = h.nc * h.nd, n = oa * ob * ncd;
    for (int idx = lane; idx < n; idx += HL_BLOCK) {
      const int ab = idx / ncd, rem = idx - ab * ncd, c = rem / h.nd, d = rem - c * h.nd;
      const int cx = c_lx(h.lc, c), cy = c_ly(h.lc, c), cz = c_lz(h.lc, c);
      const int dx = c_lx(h.ld, d), dy = c_ly(h.ld, d), dz = c_lz(h.ld, d);
      double v = 0.0;
      for (int mx = 0; mx <= dx; ++mx)
        for (int my = 0; my <= dy; ++my)
          for (int mz = 0; mz <= dz; ++mz) {
            const double cf = (double)(binom(dx, mx) * binom(dy, my) * binom(dz, mz)) * hl_pow(CD[0], dx - mx) *
                              hl_pow(CD[1], dy - my) * hl_pow(CD[2], dz - mz);
            v += cf * cur[ab * h.nf + hl_eidx(h.lc, cx + mx, cy + my, cz + mz)];
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
__global__ __launch_bounds__(HL_BLOCK) void k_jkh(JKParams prm, WaveWork ww, HLDims h) {
  extern __shared__ double lds[];
  const long long wave = blockIdx.x;
  if (wave >= ww.nwaves) return;  // block-uniform
  int lo = 0, hi = ww.nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
  }
  const int seg = __builtin_amdgcn_readfirstlane(lo);
  const int lane = threadIdx.x;
  const int k = (int)(wave - ww.choff[seg]) * 64 + lane;
  const bool active = k < ww.seg_n[seg];
  const int ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
  const ShellPair P = prm.pairs[ip];
  int iq = 0;
  bool skip = true;
  if (active) {
    iq = ww.ket[ww.seg_k0[seg] + k];
    const ShellPair Q = prm.pairs[iq];
    const bool p_first = (P.si > Q.si) || (P.si == Q.si && P.sj >= Q.sj);
    const double qbra = p_first ? P.Q : Q.Q;
    const int n = prm.nsh;
    const double* ds = prm.dsh;
    double dm = fmax(4.0 * ds[P.si * n + P.sj], 4.0 * ds[Q.si * n + Q.sj]);
    dm = fmax(dm, fmax(ds[P.sj * n + Q.sj], ds[P.sj * n + Q.si]));
    dm = fmax(dm, fmax(ds[P.si * n + Q.sj], ds[P.si * n + Q.si]));
    skip = (qbra * prm.maxden < prm.cutoff) || (P.Q * Q.Q * dm < prm.cutoff);
  }
  unsigned long long* ctr = prm.counters + 8 * (int)(wave & (OQPH_CTR_SLOTS - 1));
  count_wave(ctr + 1, active && skip);
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
    const double* V = hl_block(h, P, Q, prm.prims, prm.rys, prm.xform, 1.0e-4 * cut, lds, nprim);
    prims_used += (unsigned long long)nprim;
    if (nprim == 0) continue;
    flops += (unsigned long long)(nprim * fprim + 12 * nout);
    double deg = 1.0;
    if (P.sa == P.sb) deg *= 0.5;
    if (Q.sa == Q.sb) deg *= 0.5;
    if (ip == jq) deg *= 0.5;
    const double cj = 2.0 * deg * prm.sc, ck = -0.5 * deg * prm.se;
    const int oc = Q.ao_a, od = Q.ao_b;
    const int OA = h.oa, OB = h.ob, OC = h.oc, OD = h.od;
    auto val = [&](int a, int b, int c, int d) {
      const double v = V[((a * OB + b) * OC + c) * OD + d];
      return fabs(v) < cut ? 0.0 : v;  // OpenQP element cutoff on the AO value
    };
    auto put = [&](int x, int y, double v) {
      v *= (x == y ? 2.0 : 1.0);
      if (v != 0.0) fatomic(F + tri_idx(x, y), v);
    };
    // J_ab and J_cd
    for (int t = lane; t < OA * OB; t += HL_BLOCK) {
      const int a = t / OB, b = t - a * OB;
      double s = 0.0;
      for (int c = 0; c < OC; ++c)
        for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dp[tri_idx(oc + c, od + d)];
      put(oa + a, ob + b, cj * s);
    }
    for (int t = lane; t < OC * OD; t += HL_BLOCK) {
      const int c = t / OD, d = t - c * OD;
      double s = 0.0;
      for (int a = 0; a < OA; ++a)
        for (int b = 0; b < OB; ++b) s += val(a, b, c, d) * Dp[tri_idx(oa + a, ob + b)];
      put(oc + c, od + d, cj * s);
    }
    // K_ac += V D_bd, K_ad += V D_bc, K_bc += V D_ad, K_bd += V D_ac
    for (int t = lane; t < OA * OC; t += HL_BLOCK) {
      const int a = t / OC, c = t - a * OC;
      double s = 0.0;
      for (int b = 0; b < OB; ++b)
        for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dp[tri_idx(ob + b, od + d)];
      put(oa + a, oc + c, ck * s);
    }
    for (int t = lane; t < OA * OD; t += HL_BLOCK) {
      const int a = t / OD, d = t - a * OD;
      double s = 0.0;
      for (int b = 0; b < OB; ++b)
        for (int c = 0; c < OC; ++c) s += val(a, b, c, d) * Dp[tri_idx(ob + b, oc + c)];
      put(oa + a, od + d, ck * s);
    }
    for (int t = lane; t < OB * OC; t += HL_BLOCK) {
      const int b = t / OC, c = t - b * OC;
      double s = 0.0;
      for (int a = 0; a < OA; ++a)
        for (int d = 0; d < OD; ++d) s += val(a, b, c, d) * Dp[tri_idx(oa + a, od + d)];
      put(ob + b, oc + c, ck * s);
    }
    for (int t = lane; t < OB * OD; t += HL_BLOCK) {
      const int b = t / OD, d = t - b * OD;
      double s = 0.0;
      for (int a = 0; a < OA; ++a)
        for (int c = 0; c < OC; ++c) s += val(a, b, c, d) * Dp[tri_idx(oa + a, oc + c)];
      put(ob + b, od + d, ck * s);
    }
  }
  if (lane == 0) {
    if (flops) atomicAdd(ctr + 3, flops);
    if (prims_used) atomicAdd(ctr + 4, prims_used);
  }
}

// --------------------------------------------------------- Schwarz bounds
// One block per pair of the list: Q_ab = sqrt(max |(ab|ab)|) over the AO block.
__global__ __launch_bounds__(HL_BLOCK) void k_schwarzh(const ShellPair* pairs, const PrimPair* prims, const int* list,
                                                       int n, const double* rys, const double* xf, double qcut,
                                                       double* Qout, HLDims h) {
  extern __shared__ double lds[];
  if ((int)blockIdx.x >= n) return;
  const int ip = list[blockIdx.x];
  const ShellPair P = pairs[ip];
  int nprim = 0;
  const double* V = hl_block(h, P, P, prims, rys, xf, qcut, lds, nprim);
  double m = 0.0;
  if (nprim) {
    const int nv = h.oa * h.ob * h.oc * h.od;
    for (int i = threadIdx.x; i < nv; i += HL_BLOCK) m = fmax(m, fabs(V[i]));
  }
  m = wave_max(m);
  if (threadIdx.x == 0) Qout[ip] = sqrt(m);
}

// --------------------------------------------------------------- gradient
// runtime form of the 2D integral with bra and ket transfer:
//   I(a,b,c,d) = sum_j sum_m C(b,j) AB^(b-j) C(d,m) CD^(d-m) G(a+j, c+m)
__device__ __forceinline__ double hl_i2d(const double* g, int nk, int a, int b, int c, int d, double ab, double cd) {
  double s = 0.0;
  for (int j = 0; j <= b; ++j) {
    double t = 0.0;
    for (int m = 0; m <= d; ++m) t += (double)binom(d, m) * hl_pow(cd, d - m) * g[(a + j) * nk + c + m];
    s += (double)binom(b, j) * hl_pow(ab, b - j) * t;
  }
  return s;
}

// Gradient wavefront tasks (coarse-screened lists): a block takes one chunk
// of kets of one bra; per quartet the lanes form the Cartesian Gamma block
// in LDS, apply the fine screen, and contract it with the derivative
// integrals of their Gamma entries; per-atom sums go through LDS.
__global__ __launch_bounds__(HL_BLOCK) void k_gradh(GradParams prm, WaveWork ww, HLDims h) {
  extern __shared__ double lds[];
  double* s_de = lds + h.offB;  // 3 natm
  for (int i = threadIdx.x; i < 3 * prm.natm; i += HL_BLOCK) s_de[i] = 0.0;
  __syncthreads();
  const long long wave = blockIdx.x;
  if (wave >= ww.nwaves) return;  // block-uniform
  int lo = 0, hi = ww.nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ww.choff[mid] <= wave) lo = mid; else hi = mid;
  }
  const int seg = __builtin_amdgcn_readfirstlane(lo);
  const int lane = threadIdx.x;
  const int k = (int)(wave - ww.choff[seg]) * 64 + lane;
  const bool active = k < ww.seg_n[seg];
  const int ip = __builtin_amdgcn_readfirstlane(ww.seg_bra[seg]);
  const ShellPair P = prm.pairs[ip];
  const int iq = active ? ww.ket[ww.seg_k0[seg] + k] : 0;
  unsigned long long* ctr = prm.counters + 8 * (int)(wave & (OQPH_CTR_SLOTS - 1));
  unsigned long long todo = __ballot(active);
  double* G = lds + h.offG;
  double* Gam = lds + h.offA;
  const int NA = h.na, NB = h.nb, NC = h.nc, ND = h.nd, NG = NA * NB * NC * ND;
  const int n = prm.ncf;
  const double* Dc = prm.Dc;
  const int oa = prm.coff[P.sa], ob = prm.coff[P.sb];
  const long long fprim = flops_grad_prim_rt(h.la, h.lb, h.lc, h.ld);
  unsigned long long ncomp = 0ull, nskip = 0ull, flops = 0ull, prims_used = 0ull;
  const int nrk = h.ni * h.nk;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    todo &= todo - 1ull;
    const int jq = __builtin_amdgcn_readfirstlane(__shfl(iq, l));
    const ShellPair Q = prm.pairs[jq];
    const int oc = prm.coff[Q.sa], od = prm.coff[Q.sb];
    __syncthreads();  // the previous quartet's readers of Gamma are done
    double gm = 0.0;
    for (int t = lane; t < NG; t += HL_BLOCK) {
      const int a = t / (NB * NC * ND), r1 = t - a * NB * NC * ND, b = r1 / (NC * ND), r2 = r1 - b * NC * ND,
                c = r2 / ND, d = r2 - c * ND;
      const double g = 4.0 * prm.cs * Dc[(oa + a) * n + ob + b] * Dc[(oc + c) * n + od + d] -
                       prm.hs * (Dc[(oa + a) * n + oc + c] * Dc[(ob + b) * n + od + d] +
                                 Dc[(oa + a) * n + od + d] * Dc[(ob + b) * n + oc + c]);
      Gam[t] = g;
      gm = fmax(gm, fabs(g));
    }
    gm = wave_max(gm);
    // fine screen on the weighted contribution (grd2.F90:348)
    if (gm * P.Q * Q.Q < 0.5 * prm.gcut) {
      ++nskip;
      continue;
    }
    ++ncomp;
    const double dab2 = gm * gm;
    double fd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    int nprim = 0;
    double AB[3], CD[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      AB[d] = P.A[d] - P.B[d];
      CD[d] = Q.A[d] - Q.B[d];
    }
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
        __syncthreads();
        hl_vrr(h, s, prm.rys, G);
        __syncthreads();
        const double a2 = 2.0 * pp.a, b2 = 2.0 * pp.b, c2 = 2.0 * qq.a;
        for (int t = lane; t < NG; t += HL_BLOCK) {
          const int a = t / (NB * NC * ND), r1 = t - a * NB * NC * ND, b = r1 / (NC * ND), r2 = r1 - b * NC * ND,
                    c = r2 / ND, dd = r2 - c * ND;
          const double g = Gam[t];
          if (g == 0.0) continue;
          int ia[3], ib[3], ic[3], id[3];
          ia[0] = c_lx(h.la, a); ia[1] = c_ly(h.la, a); ia[2] = c_lz(h.la, a);
          ib[0] = c_lx(h.lb, b); ib[1] = c_ly(h.lb, b); ib[2] = c_lz(h.lb, b);
          ic[0] = c_lx(h.lc, c); ic[1] = c_ly(h.lc, c); ic[2] = c_lz(h.lc, c);
          id[0] = c_lx(h.ld, dd); id[1] = c_ly(h.ld, dd); id[2] = c_lz(h.ld, dd);
          for (int r = 0; r < h.nr; ++r) {
            double I0[3], dA[3], dB[3], dC[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              const double* gx = G + (3 * r + x) * nrk;
              const int nk = h.nk;
              I0[x] = hl_i2d(gx, nk, ia[x], ib[x], ic[x], id[x], AB[x], CD[x]);
              double v = a2 * hl_i2d(gx, nk, ia[x] + 1, ib[x], ic[x], id[x], AB[x], CD[x]);
              if (ia[x] > 0) v -= (double)ia[x] * hl_i2d(gx, nk, ia[x] - 1, ib[x], ic[x], id[x], AB[x], CD[x]);
              dA[x] = v;
              v = b2 * hl_i2d(gx, nk, ia[x], ib[x] + 1, ic[x], id[x], AB[x], CD[x]);
              if (ib[x] > 0) v -= (double)ib[x] * hl_i2d(gx, nk, ia[x], ib[x] - 1, ic[x], id[x], AB[x], CD[x]);
              dB[x] = v;
              v = c2 * hl_i2d(gx, nk, ia[x], ib[x], ic[x] + 1, id[x], AB[x], CD[x]);
              if (ic[x] > 0) v -= (double)ic[x] * hl_i2d(gx, nk, ia[x], ib[x], ic[x] - 1, id[x], AB[x], CD[x]);
              dC[x] = v;
            }
            fd[0][0] += g * dA[0] * I0[1] * I0[2];
            fd[0][1] += g * I0[0] * dA[1] * I0[2];
            fd[0][2] += g * I0[0] * I0[1] * dA[2];
            fd[1][0] += g * dB[0] * I0[1] * I0[2];
            fd[1][1] += g * I0[0] * dB[1] * I0[2];
            fd[1][2] += g * I0[0] * I0[1] * dB[2];
            fd[2][0] += g * dC[0] * I0[1] * I0[2];
            fd[2][1] += g * I0[0] * dC[1] * I0[2];
            fd[2][2] += g * I0[0] * I0[1] * dC[2];
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
    const int at[4] = {prm.shell_atom[P.sa], prm.shell_atom[P.sb], prm.shell_atom[Q.sa], prm.shell_atom[Q.sb]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double fa = wave_sum(fd[0][d]), fb = wave_sum(fd[1][d]), fc = wave_sum(fd[2][d]);
      if (lane == 0) {
        atomicAdd(&s_de[3 * at[0] + d], deg * fa);
        atomicAdd(&s_de[3 * at[1] + d], deg * fb);
        atomicAdd(&s_de[3 * at[2] + d], deg * fc);
        atomicAdd(&s_de[3 * at[3] + d], -deg * (fa + fb + fc));
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

}  // namespace oqph
