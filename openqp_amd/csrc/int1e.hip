// int1e.hip -- host one-electron integrals S, T, V for the SCF host loop and
// for building S-orthonormal test densities (OpenMP, FP64).
//
// Overlap and kinetic energy by the Obara-Saika one-dimensional recursion;
// nuclear attraction by Rys quadrature on the same root table the device
// kernels use: <a|1/r_C|b> = 2 pi/g K_ab sum_r w_r prod_d G_d(t_r^2), with
// G(i+1) = (PA - t^2 PC) G(i) + i (1 - t^2)/(2g) G(i-1) and the horizontal
// transfer (a, b+1) = (a+1, b) + AB (a, b).  Output in OpenQP's AO
// representation (component order, shells_pnrm2, pure projection) as the
// full nbf x nbf matrices OpenQP's int1 builds (source/integrals/int1.F90:147).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/oqph.h"
#include "ao_xform.h"

namespace oqph {
const double* rys_table_host(size_t* n);
const double* rys_herm_host(size_t* n);
}

namespace {

constexpr int kDeg = 12;
// root-table layout mirrored from rys_table.inc (metadata only)
#include "rys_table.inc"

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int canon(int lx, int ly, int lz) { return (ly + lz) * (ly + lz + 1) / 2 + lz; }
inline void comp(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = k - z;
}

void rys(int n, double T, const double* tab, const double* herm, double* x, double* w) {
  if (T >= RYS_TASYM[n]) {
    for (int i = 0; i < n; ++i) {
      x[i] = herm[RYS_HOFF[n] + i] / T;
      w[i] = herm[RYS_HOFF[n] + n + i] / std::sqrt(T);
    }
    return;
  }
  const int k = (int)T;
  const double s = 2.0 * (T - k) - 1.0;
  const double* c = tab + RYS_OFF[n] + (size_t)k * (2 * n * (kDeg + 1));
  for (int f = 0; f < 2 * n; ++f) {
    double v = c[f * (kDeg + 1) + kDeg];
    for (int j = kDeg - 1; j >= 0; --j) v = v * s + c[f * (kDeg + 1) + j];
    if (f < n) x[f] = v; else w[f - n] = v;
  }
}

// transform a canonical Cartesian (na x nb) block to the AO representation
void to_ao(int la, int pa, int lb, int pb, const double* xf, const double* in, double* out) {
  auto mat = [&](int l, int p, std::vector<double>& m, int& no) {
    const int nc = ncart(l);
    if (l < 2) {
      no = nc;
      m.assign(nc * nc, 0.0);
      for (int i = 0; i < nc; ++i) m[i * nc + i] = 1.0;
      return;
    }
    no = (p && l >= 2) ? 2 * l + 1 : nc;
    m.assign(no * nc, 0.0);
    for (int o = 0; o < no; ++o)
      for (int i = 0; i < nc; ++i) m[o * nc + i] = xf[(((l - 2) * 2 + (p ? 1 : 0)) * 10 + o) * 10 + i];
  };
  std::vector<double> A, B;
  int oa, ob;
  mat(la, pa, A, oa);
  mat(lb, pb, B, ob);
  const int na = ncart(la), nb = ncart(lb);
  for (int i = 0; i < oa; ++i)
    for (int j = 0; j < ob; ++j) {
      double s = 0.0;
      for (int p = 0; p < na; ++p) {
        const double ap = A[i * na + p];
        if (ap == 0.0) continue;
        for (int q = 0; q < nb; ++q) s += ap * B[j * nb + q] * in[p * nb + q];
      }
      out[i * ob + j] = s;
    }
}

}  // namespace

extern "C" int oqph_int1e(int nsh, const int* am, const int* harmonic, const int* atom, const int* nprim,
                          const double* ex, const double* cc, int natm, const double* xyz, const double* zn,
                          double* S, double* T, double* V) {
  if (nsh <= 0 || !am || !harmonic || !atom || !nprim || !ex || !cc || natm <= 0 || !xyz || !zn) return 4;
  size_t nt, nh;
  const double* tab = oqph::rys_table_host(&nt);
  const double* herm = oqph::rys_herm_host(&nh);
  double xf[2][2][10][10];
  oqph_host::build_xform(xf);
  std::vector<int> poff(nsh), aoff(nsh), nao(nsh);
  int p = 0, a = 0;
  for (int i = 0; i < nsh; ++i) {
    if (am[i] < 0 || am[i] > 3 || nprim[i] <= 0 || atom[i] < 0 || atom[i] >= natm) return 4;
    poff[i] = p;
    aoff[i] = a;
    nao[i] = (harmonic[i] && am[i] >= 2) ? 2 * am[i] + 1 : ncart(am[i]);
    p += nprim[i];
    a += nao[i];
  }
  const int nbf = a;
#pragma omp parallel for schedule(dynamic, 1)
  for (int sa = 0; sa < nsh; ++sa) {
    std::vector<double> bs(100), bt(100), bv(100), out(100);
    for (int sb = 0; sb < nsh; ++sb) {
      const int la = am[sa], lb = am[sb], na = ncart(la), nb = ncart(lb), L = la + lb;
      const double* Ac = xyz + 3 * atom[sa];
      const double* Bc = xyz + 3 * atom[sb];
      double AB[3], ab2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        AB[d] = Ac[d] - Bc[d];
        ab2 += AB[d] * AB[d];
      }
      std::fill(bs.begin(), bs.end(), 0.0);
      std::fill(bt.begin(), bt.end(), 0.0);
      std::fill(bv.begin(), bv.end(), 0.0);
      for (int ip = 0; ip < nprim[sa]; ++ip)
        for (int iq = 0; iq < nprim[sb]; ++iq) {
          const double al = ex[poff[sa] + ip], be = ex[poff[sb] + iq], g = al + be;
          const double K = cc[poff[sa] + ip] * cc[poff[sb] + iq] * std::exp(-al * be / g * ab2);
          double P[3], PA[3], PB[3];
          for (int d = 0; d < 3; ++d) {
            P[d] = (al * Ac[d] + be * Bc[d]) / g;
            PA[d] = P[d] - Ac[d];
            PB[d] = P[d] - Bc[d];
          }
          // 1D overlaps s[d][i][j], i <= la, j <= lb + 2
          double s1[3][5][7];
          const double h = 0.5 / g;
          for (int d = 0; d < 3; ++d) {
            for (int i = 0; i <= la; ++i)
              for (int j = 0; j <= lb + 2; ++j) {
                double v;
                if (i == 0 && j == 0) v = 1.0;
                else if (i > 0) {
                  v = PA[d] * s1[d][i - 1][j];
                  if (i > 1) v += (i - 1) * h * s1[d][i - 2][j];
                  if (j > 0) v += j * h * s1[d][i - 1][j - 1];
                } else {
                  v = PB[d] * s1[d][0][j - 1];
                  if (j > 1) v += (j - 1) * h * s1[d][0][j - 2];
                }
                s1[d][i][j] = v;
              }
          }
          const double s0 = K * std::pow(M_PI / g, 1.5);
          for (int ia = 0; ia < na; ++ia)
            for (int ib = 0; ib < nb; ++ib) {
              int ax, ay, az, bx, by, bz;
              comp(la, ia, ax, ay, az);
              comp(lb, ib, bx, by, bz);
              const int ai[3] = {ax, ay, az}, bi[3] = {bx, by, bz};
              double sv[3], tv[3];
              for (int d = 0; d < 3; ++d) {
                const int i = ai[d], j = bi[d];
                sv[d] = s1[d][i][j];
                double t = -2.0 * be * be * s1[d][i][j + 2] + be * (2 * j + 1) * s1[d][i][j];
                if (j >= 2) t -= 0.5 * j * (j - 1) * s1[d][i][j - 2];
                tv[d] = t;
              }
              bs[ia * nb + ib] += s0 * sv[0] * sv[1] * sv[2];
              bt[ia * nb + ib] += s0 * (tv[0] * sv[1] * sv[2] + sv[0] * tv[1] * sv[2] + sv[0] * sv[1] * tv[2]);
            }
          // nuclear attraction (Rys)
          const int nr = L / 2 + 1;
          for (int c = 0; c < natm; ++c) {
            if (zn[c] == 0.0) continue;
            const double* Cc = xyz + 3 * c;
            double PC[3], pc2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              PC[d] = P[d] - Cc[d];
              pc2 += PC[d] * PC[d];
            }
            double rx[8], rw[8];
            rys(nr, g * pc2, tab, herm, rx, rw);
            const double pre = -zn[c] * K * 2.0 * M_PI / g;
            for (int r = 0; r < nr; ++r) {
              const double t2 = rx[r];
              double G[3][7];  // bra-built 1D integrals up to L, then HRR per component
              for (int d = 0; d < 3; ++d) {
                const double c00 = PA[d] - t2 * PC[d], b10 = (1.0 - t2) * h;
                G[d][0] = 1.0;
                if (L >= 1) G[d][1] = c00;
                for (int i = 1; i < L; ++i) G[d][i + 1] = c00 * G[d][i] + i * b10 * G[d][i - 1];
              }
              for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nb; ++ib) {
                  int ax, ay, az, bx, by, bz;
                  comp(la, ia, ax, ay, az);
                  comp(lb, ib, bx, by, bz);
                  const int ai[3] = {ax, ay, az}, bi[3] = {bx, by, bz};
                  double prod = pre * rw[r];
                  for (int d = 0; d < 3; ++d) {
                    // (a, b) = sum_k C(b,k) AB^(b-k) (a+k, 0)
                    double v = 0.0, pw = 1.0;
                    for (int k = bi[d]; k >= 0; --k) {
                      double bin = 1.0;
                      for (int q = 1; q <= k; ++q) bin = bin * (bi[d] - q + 1) / q;
                      v += bin * pw * G[d][ai[d] + k];
                      pw *= AB[d];
                    }
                    prod *= v;
                  }
                  bv[ia * nb + ib] += prod;
                }
            }
          }
        }
      double* dst[3] = {S, T, V};
      double* src[3] = {bs.data(), bt.data(), bv.data()};
      for (int w = 0; w < 3; ++w) {
        if (!dst[w]) continue;
        to_ao(la, harmonic[sa], lb, harmonic[sb], &xf[0][0][0][0], src[w], out.data());
        for (int i = 0; i < nao[sa]; ++i)
          for (int j = 0; j < nao[sb]; ++j) dst[w][(size_t)(aoff[sa] + i) * nbf + aoff[sb] + j] = out[i * nao[sb] + j];
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// One-electron part of the RHF nuclear gradient (host, OpenMP, FP64):
//   grad += sum_mn D_mn d(T + V)_mn/dR - sum_mn W_mn dS_mn/dR
// with D the total density and W the energy-weighted density
// W = 2 C_occ diag(eps_occ) C_occ^T, both full nbf x nbf in the AO order of
// oqph_int1e.  Covers OpenQP's grad_ee_overlap, grad_ee_kinetic,
// grad_en_hellman_feynman and grad_en_pulay (source/integrals/grd1.F90,
// driver source/modules/hf_gradient.F90:143-224; W as eijden builds it,
// grd1.F90:127-195, with the sign moved into the overlap term).  The nuclear
// repulsion term is the caller's.
//
// Derivatives of a Cartesian Gaussian: d/dA_x x_A^a e^{-alpha r_A^2} =
// 2 alpha x_A^{a+1} ... - a x_A^{a-1} ...; two-centre integrals are
// translationally invariant (d/dB = -d/dA); for the nuclear attraction of
// nucleus C, d/dC = -(d/dA + d/dB) of that term.  The AO-basis densities are
// contracted in the Cartesian basis of each shell pair, Dc = X_a^T D X_b.
extern "C" int oqph_grad1e(int nsh, const int* am, const int* harmonic, const int* atom, const int* nprim,
                           const double* ex, const double* cc, int natm, const double* xyz, const double* zn,
                           const double* D, const double* W, double* grad) {
  if (nsh <= 0 || !am || !harmonic || !atom || !nprim || !ex || !cc || natm <= 0 || !xyz || !zn || !D || !W ||
      !grad)
    return 4;
  size_t nt, nh;
  const double* tab = oqph::rys_table_host(&nt);
  const double* herm = oqph::rys_herm_host(&nh);
  double xf[2][2][10][10];
  oqph_host::build_xform(xf);
  std::vector<int> poff(nsh), aoff(nsh), nao(nsh);
  int p = 0, a = 0;
  for (int i = 0; i < nsh; ++i) {
    if (am[i] < 0 || am[i] > 3 || nprim[i] <= 0 || atom[i] < 0 || atom[i] >= natm) return 4;
    poff[i] = p;
    aoff[i] = a;
    nao[i] = (harmonic[i] && am[i] >= 2) ? 2 * am[i] + 1 : ncart(am[i]);
    p += nprim[i];
    a += nao[i];
  }
  const int nbf = a;
  // AO transform of one shell: X (no x nc), identity for l < 2
  auto xmat = [&](int l, int pure, std::vector<double>& m) {
    const int nc = ncart(l);
    const int no = (pure && l >= 2) ? 2 * l + 1 : nc;
    m.assign((size_t)no * nc, 0.0);
    if (l < 2) {
      for (int i = 0; i < nc; ++i) m[i * nc + i] = 1.0;
      return;
    }
    for (int o = 0; o < no; ++o)
      for (int i = 0; i < nc; ++i) m[o * nc + i] = xf[l - 2][pure ? 1 : 0][o][i];
  };
  std::vector<double> acc((size_t)3 * natm, 0.0);
#pragma omp parallel
  {
    std::vector<double> g((size_t)3 * natm, 0.0);
    std::vector<double> Xa, Xb, Dc, Wc;
#pragma omp for schedule(dynamic, 1)
    for (int sa = 0; sa < nsh; ++sa) {
      for (int sb = 0; sb <= sa; ++sb) {
        const int la = am[sa], lb = am[sb], na = ncart(la), nb = ncart(lb), L = la + lb;
        const int oa = nao[sa], ob = nao[sb];
        const double fac = (sa == sb) ? 1.0 : 2.0;  // the (sb, sa) block contributes the same
        xmat(la, harmonic[sa], Xa);
        xmat(lb, harmonic[sb], Xb);
        Dc.assign((size_t)na * nb, 0.0);
        Wc.assign((size_t)na * nb, 0.0);
        for (int i = 0; i < oa; ++i)
          for (int j = 0; j < ob; ++j) {
            const size_t ij = (size_t)(aoff[sa] + i) * nbf + aoff[sb] + j;
            const double dv = D[ij], wv = W[ij];
            if (dv == 0.0 && wv == 0.0) continue;
            for (int pp = 0; pp < na; ++pp) {
              const double xa = Xa[i * na + pp];
              if (xa == 0.0) continue;
              for (int q = 0; q < nb; ++q) {
                const double x = xa * Xb[j * nb + q];
                Dc[pp * nb + q] += x * dv;
                Wc[pp * nb + q] += x * wv;
              }
            }
          }
        const int ca = atom[sa], cb = atom[sb];
        const double* Ac = xyz + 3 * ca;
        const double* Bc = xyz + 3 * cb;
        double AB[3], ab2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          AB[d] = Ac[d] - Bc[d];
          ab2 += AB[d] * AB[d];
        }
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};  // basis-centre derivatives (T, V) and -W dS
        for (int ip = 0; ip < nprim[sa]; ++ip)
          for (int iq = 0; iq < nprim[sb]; ++iq) {
            const double al = ex[poff[sa] + ip], be = ex[poff[sb] + iq], gg = al + be;
            const double K = cc[poff[sa] + ip] * cc[poff[sb] + iq] * std::exp(-al * be / gg * ab2);
            double P[3], PA[3], PB[3];
            for (int d = 0; d < 3; ++d) {
              P[d] = (al * Ac[d] + be * Bc[d]) / gg;
              PA[d] = P[d] - Ac[d];
              PB[d] = P[d] - Bc[d];
            }
            const double h = 0.5 / gg;
            // 1D overlaps s1[d][i][j], i <= la + 1, j <= lb + 2
            double s1[3][6][7];
            for (int d = 0; d < 3; ++d)
              for (int i = 0; i <= la + 1; ++i)
                for (int j = 0; j <= lb + 2; ++j) {
                  double v;
                  if (i == 0 && j == 0) v = 1.0;
                  else if (i > 0) {
                    v = PA[d] * s1[d][i - 1][j];
                    if (i > 1) v += (i - 1) * h * s1[d][i - 2][j];
                    if (j > 0) v += j * h * s1[d][i - 1][j - 1];
                  } else {
                    v = PB[d] * s1[d][0][j - 1];
                    if (j > 1) v += (j - 1) * h * s1[d][0][j - 2];
                  }
                  s1[d][i][j] = v;
                }
            auto t1 = [&](int d, int i, int j) {  // 1D kinetic on the ket
              double t = -2.0 * be * be * s1[d][i][j + 2] + be * (2 * j + 1) * s1[d][i][j];
              if (j >= 2) t -= 0.5 * j * (j - 1) * s1[d][i][j - 2];
              return t;
            };
            const double s0 = K * std::pow(M_PI / gg, 1.5);
            for (int ia = 0; ia < na; ++ia)
              for (int ib = 0; ib < nb; ++ib) {
                const double dcv = Dc[ia * nb + ib], wcv = Wc[ia * nb + ib];
                if (dcv == 0.0 && wcv == 0.0) continue;
                int ax, ay, az, bx, by, bz;
                comp(la, ia, ax, ay, az);
                comp(lb, ib, bx, by, bz);
                const int ai[3] = {ax, ay, az}, bi[3] = {bx, by, bz};
                double sv[3], tv[3], ds[3], dt[3];
                for (int d = 0; d < 3; ++d) {
                  const int i = ai[d], j = bi[d];
                  sv[d] = s1[d][i][j];
                  tv[d] = t1(d, i, j);
                  ds[d] = 2.0 * al * s1[d][i + 1][j] - (i > 0 ? i * s1[d][i - 1][j] : 0.0);
                  dt[d] = 2.0 * al * t1(d, i + 1, j) - (i > 0 ? i * t1(d, i - 1, j) : 0.0);
                }
                for (int d = 0; d < 3; ++d) {
                  const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                  const double dS = s0 * ds[d] * sv[e1] * sv[e2];
                  const double dT = s0 * (dt[d] * sv[e1] * sv[e2] + ds[d] * (tv[e1] * sv[e2] + sv[e1] * tv[e2]));
                  const double v = dcv * dT - wcv * dS;
                  gA[d] += v;
                  gB[d] -= v;
                }
              }
            // nuclear attraction: derivatives on A and B per nucleus C, C by invariance
            const int nr = (L + 1) / 2 + 1;
            for (int c = 0; c < natm; ++c) {
              if (zn[c] == 0.0) continue;
              const double* Cc = xyz + 3 * c;
              double PC[3], pc2 = 0.0;
              for (int d = 0; d < 3; ++d) {
                PC[d] = P[d] - Cc[d];
                pc2 += PC[d] * PC[d];
              }
              double rx[8], rw[8];
              rys(nr, gg * pc2, tab, herm, rx, rw);
              const double pre = -zn[c] * K * 2.0 * M_PI / gg;
              double vA[3] = {0, 0, 0}, vB[3] = {0, 0, 0};
              for (int r = 0; r < nr; ++r) {
                const double t2 = rx[r];
                double G[3][9], I[3][6][6];
                for (int d = 0; d < 3; ++d) {
                  const double c00 = PA[d] - t2 * PC[d], b10 = (1.0 - t2) * h;
                  G[d][0] = 1.0;
                  if (L + 1 >= 1) G[d][1] = c00;
                  for (int i = 1; i < L + 1; ++i) G[d][i + 1] = c00 * G[d][i] + i * b10 * G[d][i - 1];
                  // I(i, j) = sum_k C(j,k) AB^(j-k) G(i+k), i <= la+1, j <= lb+1, i + j <= L + 1
                  for (int i = 0; i <= la + 1; ++i)
                    for (int j = 0; j <= lb + 1 && i + j <= L + 1; ++j) {
                      double v = 0.0, pw = 1.0;
                      for (int k = j; k >= 0; --k) {
                        double bin = 1.0;
                        for (int q = 1; q <= k; ++q) bin = bin * (j - q + 1) / q;
                        v += bin * pw * G[d][i + k];
                        pw *= AB[d];
                      }
                      I[d][i][j] = v;
                    }
                }
                const double wr = pre * rw[r];
                for (int ia = 0; ia < na; ++ia)
                  for (int ib = 0; ib < nb; ++ib) {
                    const double dcv = Dc[ia * nb + ib];
                    if (dcv == 0.0) continue;
                    int ax, ay, az, bx, by, bz;
                    comp(la, ia, ax, ay, az);
                    comp(lb, ib, bx, by, bz);
                    const int ai[3] = {ax, ay, az}, bi[3] = {bx, by, bz};
                    double val[3], dA[3], dB[3];
                    for (int d = 0; d < 3; ++d) {
                      const int i = ai[d], j = bi[d];
                      val[d] = I[d][i][j];
                      dA[d] = 2.0 * al * I[d][i + 1][j] - (i > 0 ? i * I[d][i - 1][j] : 0.0);
                      dB[d] = 2.0 * be * I[d][i][j + 1] - (j > 0 ? j * I[d][i][j - 1] : 0.0);
                    }
                    for (int d = 0; d < 3; ++d) {
                      const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                      vA[d] += wr * dcv * dA[d] * val[e1] * val[e2];
                      vB[d] += wr * dcv * dB[d] * val[e1] * val[e2];
                    }
                  }
              }
              for (int d = 0; d < 3; ++d) {
                gA[d] += vA[d];
                gB[d] += vB[d];
                g[3 * c + d] -= fac * (vA[d] + vB[d]);
              }
            }
          }
        for (int d = 0; d < 3; ++d) {
          g[3 * ca + d] += fac * gA[d];
          g[3 * cb + d] += fac * gB[d];
        }
      }
    }
#pragma omp critical
    for (size_t i = 0; i < g.size(); ++i) acc[i] += g[i];
  }
  for (size_t i = 0; i < acc.size(); ++i) grad[i] += acc[i];
  return 0;
}
