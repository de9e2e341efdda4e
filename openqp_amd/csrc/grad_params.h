// grad_params.h -- launch arguments of the 2e-gradient kernels (host and device).
#pragma once
#include "oqph_common.h"

namespace oqph {

// The per-block atomic flush of the atomic gradient goes to one of
// OQPH_DE_SLOTS copies of the natm x 3 array (block index mod slots), summed
// by the host, so thousands of blocks do not serialise on the same lines.
#define OQPH_DE_SLOTS 32

struct GradParams {
  const ShellPair* pairs;
  const PrimPair* prims;
  const double* rys;
  const float* rys32;     // FP32 root fits (the FP32 form of k_grad)
  const double* Dc;       // ncf x ncf raw Cartesian effective density
  const int* coff;        // Cartesian offset per shell
  const int* shell_atom;  // atom of each shell
  double* de;             // OQPH_DE_SLOTS x natm*3
  unsigned long long* counters;  // [0] computed, [1] fine-screened, [3] model flops, [4] prim quartets
  int ncf, natm;
  double gcut;            // grad_cutoff
  double dabcut2;         // (two-particle density cutoff)^2
  double cs, hs;          // Coulomb and exchange scales
  // MRSF two-particle density (the MRSF forms of k_grad and k_gradh): when Mc is set,
  // Gamma comes from the eleven Cartesian matrices Mc[x ncf^2 ..] (row-major,
  // M(i, j) = Mc[x ncf^2 + i ncf + j]): d2a, d2b, p2a, p2b, ball, bo2v, bo1v,
  // bco1, bco2, o21v, co12, with the response exchange scale hs2, the
  // spin-pair scales qf[3] and the multiplicity sign sgnk; Dc is unused.
  const double* Mc;
  double hs2, qf0, qf1, qf2, sgnk;
};

// MRSF two-particle density of Cartesian components (i j | k l)
// (grd2_mrsf_compute_data_t_get_density, tdhf_mrsf_gradient.F90:449-596):
// the ground-state Coulomb and exchange parts of the spin sums and
// differences with their relaxed difference densities, the response ball
// exchange, and the spin-pair products of the response densities.
__device__ inline double mrsf_gamma(const GradParams& prm, int i, int j, int k, int l) {
  const size_t n = prm.ncf, n2 = n * n;
  const double* m = prm.Mc;
  auto M = [&](int x, int r, int c) { return m[x * n2 + (size_t)r * n + c]; };
  enum { D2A = 0, D2B, P2A, P2B, BALL, BO2V, BO1V, BCO1, BCO2, O21V, CO12 };
  double g = 4.0 * prm.cs * ((M(D2A, i, j) + M(P2A, i, j)) * M(D2A, k, l) + M(D2A, i, j) * M(P2A, k, l));
  if (prm.hs != 0.0 || prm.hs2 != 0.0) {
    const double dq = (M(D2A, i, k) + M(P2A, i, k)) * M(D2A, j, l) + M(D2A, i, k) * M(P2A, j, l) +
                      (M(D2A, i, l) + M(P2A, i, l)) * M(D2A, j, k) + M(D2A, i, l) * M(P2A, j, k) +
                      (M(D2B, i, k) + M(P2B, i, k)) * M(D2B, j, l) + M(D2B, i, k) * M(P2B, j, l) +
                      (M(D2B, i, l) + M(P2B, i, l)) * M(D2B, j, k) + M(D2B, i, l) * M(P2B, j, k);
    const double dt = M(BALL, i, k) * M(BALL, j, l) + M(BALL, k, i) * M(BALL, l, j) +
                      M(BALL, i, l) * M(BALL, j, k) + M(BALL, l, i) * M(BALL, k, j);
    g -= prm.hs * dq + 2.0 * prm.hs2 * dt;
  }
  // x(ik) y(jl) + x(il) y(jk) + x(jk) y(il) + x(jl) y(ik) and the same with
  // both factors transposed; x = y: the co12 / o21v couplings
  auto cross = [&](int x, int y) {
    return M(x, i, k) * M(y, j, l) + M(x, i, l) * M(y, j, k) + M(x, j, k) * M(y, i, l) + M(x, j, l) * M(y, i, k) +
           M(x, l, j) * M(y, k, i) + M(x, k, j) * M(y, l, i) + M(x, l, i) * M(y, k, j) + M(x, k, i) * M(y, l, j);
  };
  auto direct = [&](int x, int y) {
    const double xij = M(x, i, j) + M(x, j, i), xkl = M(x, k, l) + M(x, l, k);
    const double yij = M(y, i, j) + M(y, j, i), ykl = M(y, k, l) + M(y, l, k);
    return xij * ykl + xkl * yij;
  };
  if (prm.qf0 != 0.0) {  // co12 x co12: x(ik) x(lj) + ..., each product twice
    const double b = M(CO12, i, k) * M(CO12, l, j) + M(CO12, i, l) * M(CO12, k, j) +
                     M(CO12, j, k) * M(CO12, l, i) + M(CO12, j, l) * M(CO12, k, i);
    g += prm.sgnk * prm.qf0 * 2.0 * b;
  }
  if (prm.qf1 != 0.0) {
    const double b = M(O21V, i, k) * M(O21V, l, j) + M(O21V, i, l) * M(O21V, k, j) +
                     M(O21V, j, k) * M(O21V, l, i) + M(O21V, j, l) * M(O21V, k, i);
    g += prm.sgnk * prm.qf1 * 2.0 * b;
  }
  if (prm.qf2 != 0.0)
    g += prm.sgnk * prm.qf2 *
         (-cross(BCO1, BO2V) - cross(BCO2, BO1V) - cross(BO2V, BCO1) - cross(BO1V, BCO2) + direct(BCO1, BO2V) +
          direct(BCO2, BO1V) + direct(BO2V, BCO1) + direct(BO1V, BCO2));
  return g;
}

}  // namespace oqph
