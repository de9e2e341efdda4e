// ao_xform.h -- canonical-Cartesian -> OpenQP AO transforms (host).
//
// Canonical order inside a shell: lx descending, then ly descending.  OpenQP
// AO order and component norms: source/constants.F90:34-60, 110+; pure
// components m = -l..l are the real solid harmonics on unit-normalised
// Cartesians (Helgaker, Jorgensen, Olsen, eqs. 6.4.47-6.4.50), orthonormal in
// the intra-shell metric as source/integrals/cart2sph.F90 requires.
#pragma once
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace oqph_host {
// real solid harmonics on unit-normalised Cartesians, m = -l..l (Helgaker,
// Jorgensen, Olsen, eqs. 6.4.47-6.4.50), orthonormal in the intra-shell
// metric; OpenQP component order (constants.F90:34-60) and norms (:110+).
static const int ORD2[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
static const int ORD3[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1},
                                {1, 2, 0}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {1, 1, 1}};
inline double dfac(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}
inline double binomd(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - i + 1) / i;
  return r;
}
inline void build_xform(double out[2][2][10][10]) {
  memset(out, 0, sizeof(double) * 400);
  for (int l = 2; l <= 3; ++l) {
    const int(*ord)[3] = (l == 2) ? ORD2 : ORD3;
    const int nc = (l + 1) * (l + 2) / 2;
    // canonical local index of a component (lx desc, ly desc)
    auto canon = [&](int lx, int ly, int lz) { return (ly + lz) * (ly + lz + 1) / 2 + lz; };
    // Cartesian: AO row q <- canonical col canon(q) * pnrm2
    for (int q = 0; q < nc; ++q) {
      const int* c = ord[q];
      double nrm = std::sqrt(dfac(2 * l - 1) / (dfac(2 * c[0] - 1) * dfac(2 * c[1] - 1) * dfac(2 * c[2] - 1)));
      out[l - 2][0][q][canon(c[0], c[1], c[2])] = nrm;
    }
    // pure: row s (m = s - l) <- sum_q c2s[q][s] * pnrm2[q] on canonical col
    for (int m = -l; m <= l; ++m) {
      const int am = std::abs(m);
      double row[10] = {0};
      const int k0 = m >= 0 ? 0 : 1;
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u)
          for (int k = k0; k <= am; k += 2) {
            const int sp = t + (k - k0) / 2;
            double c = ((sp & 1) ? -1.0 : 1.0) * std::pow(0.25, t) * binomd(l, t) * binomd(l - t, am + t) *
                       binomd(t, u) * binomd(am, k);
            if (c == 0.0) continue;
            const int ax = 2 * t + am - 2 * u - k, ay = 2 * u + k, az = l - 2 * t - am;
            if (ax < 0 || ay < 0 || az < 0) continue;
            int ci = -1;
            for (int q = 0; q < nc; ++q)
              if (ord[q][0] == ax && ord[q][1] == ay && ord[q][2] == az) ci = q;
            row[ci] += c * std::sqrt(dfac(2 * ax - 1) * dfac(2 * ay - 1) * dfac(2 * az - 1));
          }
      double n2 = 0.0;
      for (int p = 0; p < nc; ++p)
        for (int q = 0; q < nc; ++q) {
          double s = 1.0, sp_ = 1.0, sq_ = 1.0;
          for (int d = 0; d < 3; ++d) {
            int e = ord[p][d] + ord[q][d];
            s = (e & 1) ? 0.0 : s * dfac(e - 1);
            sp_ *= dfac(2 * ord[p][d] - 1);
            sq_ *= dfac(2 * ord[q][d] - 1);
          }
          n2 += row[p] * row[q] * s / std::sqrt(sp_ * sq_);
        }
      for (int q = 0; q < nc; ++q) {
        const int* c = ord[q];
        double nrm = std::sqrt(dfac(2 * l - 1) / (dfac(2 * c[0] - 1) * dfac(2 * c[1] - 1) * dfac(2 * c[2] - 1)));
        out[l - 2][1][m + l][canon(c[0], c[1], c[2])] += row[q] / std::sqrt(n2) * nrm;
      }
    }
  }
}

}  // namespace oqph_host
