"""Write the committed golden J/K fixtures from the CPU oracle.

Fixtures: tests/golden/<name>.npz holding the packed density d, the packed
2e Fock f (fock_jk semantics: sc J - se K/2, se = sc = 1, int2e_cutoff
5e-11) and the Schwarz table, for
  h2o8_631g      (H2O)8 / 6-31G, seeded orientations, 40 occupied (104 bf);
  h2o_631gs      H2O / 6-31G* Cartesian d at the OpenQP KAT geometry, 5 occupied;
  h2o2_ccpvdz    (H2O)2 / cc-pVDZ pure d, 10 occupied;
  h2o_ccpvtz     H2O / cc-pVTZ pure d and f at the OpenQP KAT geometry, 5 occupied.
Run from the repo root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from helpers import workload, small_system, density_for  # noqa: E402


def make(name, bas, nocc, seed=0):
    D, dp = density_for(O, bas, nocc, seed)
    Q = O.schwarz(bas)
    f, st = O.fock_jk(bas, dp, Q=Q)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), d=dp, f=f, Q=Q, stats=st)
    print(name, bas.nbf, st)


if __name__ == "__main__":
    b, nocc = workload("h2o8_631g")
    make("h2o8_631g", b, nocc)
    make("h2o_631gs", small_system("6-31g(d)"), 5)
    make("h2o2_ccpvdz", small_system("cc-pvdz", nw=2), 10)
    make("h2o_ccpvtz", small_system("cc-pvtz"), 5)
