"""Golden fixture for the C20H42 / 6-31G(d) 2e gradient (the analytic-gradient
workload): the oracle's dE2/dR on all 62 atoms (186 coordinates) at a seeded
S-orthonormal density (81 occupied orbitals, seed 0, D = 2 C C^T), written
with the screening thresholds the device uses (grad_cutoff 1e-10).  The
density itself is not stored: the test rebuilds it from the same seed.

    python tests/golden/make_alkane_grad.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from openqp_amd import basis as B, molecules as M  # noqa: E402
from openqp_amd.lib import int1e  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    bld, bname, nocc = M.WORKLOADS["c20h42_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    S, _, _ = int1e(bas)
    dp = B.pack_tri(M.random_density(S, nocc, seed=0))
    Q = O.schwarz(bas)
    de, st = O.grad2(bas, dp, Q=Q)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c20h42_631gd_grad2.npz")
    np.savez_compressed(out, de=de, seed=0, nocc=nocc, stats=st, d_checksum=float(np.sum(dp * dp)),
                        xyz=bas.xyz)
    print(out, st)


if __name__ == "__main__":
    main()
