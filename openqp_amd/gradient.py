"""RHF nuclear gradient at a converged SCF solution.

dE/dR = dE_nn/dR + sum D dH/dR - sum W dS/dR + dE2/dR, the sum OpenQP forms
in hf_gradient (source/modules/hf_gradient.F90: hf_1e_grad :143-224 for the
nuclear-repulsion, overlap, kinetic, Hellmann-Feynman and Pulay terms,
hf_2e_grad :229-337 for the two-electron term), with

* D = 2 C_occ C_occ^T the total density and W = 2 C_occ diag(eps_occ) C_occ^T
  the energy-weighted density (eijden, source/integrals/grd1.F90:127-195,
  which stores -W and adds the overlap term with a plus sign);
* the one-electron terms from liboqph's host code (oqph_grad1e), the
  two-electron term from the device adjoint kernel (oqph_grad2).
"""
from __future__ import annotations

import numpy as np

from . import basis as B
from .lib import FockBuilder, grad1e


def nuclear_repulsion_gradient(zn: np.ndarray, xyz: np.ndarray) -> np.ndarray:
    """d/dR_A of sum_{A<B} Z_A Z_B / |R_A - R_B| (natm x 3, Eh/bohr)."""
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    zn = np.asarray(zn, np.float64)
    g = np.zeros_like(xyz)
    for a in range(len(zn)):
        for b in range(len(zn)):
            if a == b:
                continue
            r = xyz[a] - xyz[b]
            g[a] -= zn[a] * zn[b] * r / np.linalg.norm(r) ** 3
    return g


def rhf_gradient(bas: B.Basis, orbitals: np.ndarray, orbital_energies: np.ndarray, nocc: int,
                 builder: FockBuilder) -> dict:
    """Total RHF gradient (natm x 3) and its parts from converged orbitals.

    ``builder`` must hold the same basis and geometry; its thresholds set the
    two-electron screening (grad_cutoff)."""
    C = orbitals[:, :nocc]
    eps = orbital_energies[:nocc]
    D = 2.0 * C @ C.T
    W = 2.0 * (C * eps) @ C.T
    g1 = grad1e(bas, D, W)
    g2 = builder.grad2(B.pack_tri(D)).reshape(bas.natm, 3)
    gn = nuclear_repulsion_gradient(bas.zn, bas.xyz)
    return {"total": g1 + g2 + gn, "one_electron": g1, "two_electron": g2, "nuclear": gn}
