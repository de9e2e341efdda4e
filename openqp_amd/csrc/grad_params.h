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
  // MRSF two-particle density (run-time-class kernel only): when Mc is set,
  // Gamma comes from the eleven Cartesian matrices Mc[x ncf^2 ..] (row-major,
  // M(i, j) = Mc[x ncf^2 + i ncf + j]): d2a, d2b, p2a, p2b, ball, bo2v, bo1v,
  // bco1, bco2, o21v, co12, with the response exchange scale hs2, the
  // spin-pair scales qf[3] and the multiplicity sign sgnk; Dc is unused.
  const double* Mc;
  double hs2, qf0, qf1, qf2, sgnk;
};

}  // namespace oqph
