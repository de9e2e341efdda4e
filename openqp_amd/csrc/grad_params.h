// grad_params.h -- launch arguments of the 2e-gradient kernels (host and device).
#pragma once
#include "oqph_common.h"

namespace oqph {

struct GradParams {
  const ShellPair* pairs;
  const PrimPair* prims;
  const double* rys;
  const double* herm;
  const double* Dc;       // ncf x ncf raw Cartesian effective density
  const int* coff;        // Cartesian offset per shell
  const int* shell_atom;  // atom of each shell
  double* de;             // natm*3
  unsigned long long* counters;  // [0] computed, [1] fine-screened, [3] model flops, [4] prim quartets
  int ncf, natm;
  double gcut;            // grad_cutoff
  double dabcut2;         // (two-particle density cutoff)^2
  double cs, hs;          // Coulomb and exchange scales
};

}  // namespace oqph
