"""Host SCF logic on the CPU: the extended-Hueckel guess and the conditioned
DIIS, driven by an FP64 CPU Fock build (the oracle standing in for the device
builder, so the host loop is checked without a GPU).  The converged energy is
pinned by OpenQP's own known answer."""
import json
import os
import warnings

import numpy as np
import pytest

from openqp_amd import basis as B, molecules as M
from openqp_amd.guess import huckel_guess, huckel_minimal
from openqp_amd.lib import int1e
from helpers import ROOT, small_system


class _CPUBuilder:
    """fock_jk / fock_jk_inc / vxc with the oracle's FP64 builds (test stand-in)."""

    def __init__(self, oracle, bas):
        self.O, self.bas = oracle, bas
        self.Q = oracle.schwarz(bas)

    def fock_jk(self, d, se=1.0):
        return self.O.fock_jk(self.bas, d, se=se, Q=self.Q)[0]

    def fock_jk_inc(self, d, d_old, f_old, se=1.0):
        f = self.fock_jk(d - d_old, se) + f_old
        f_old[:] = f
        d_old[:] = d
        return f

    def fock_jk_uhf(self, da, db, se=1.0):
        fa, fb, _ = self.O.fock_jk_uhf(self.bas, da, db, se=se, Q=self.Q)
        return fa, fb

    def set_xc(self, name, nrad, ntheta):
        from oracle import xc as X
        self.xc, self.X = name, X
        pts, w, _ = X.molecular_grid(self.bas.zn, self.bas.xyz, nrad, ntheta)
        self.grid = (pts, w)
        self.ao = X.ao_values(self.bas, pts, self.O)

    def vxc(self, dp, nfocks=1):
        n = dp.size // nfocks
        ds = [B.unpack_tri(dp[i * n:(i + 1) * n], self.bas.nbf) for i in range(nfocks)]
        V, exc, nel = self.X.vxc(self.bas, ds[0] if nfocks == 1 else ds, self.xc, self.grid, self.O, nfocks, self.ao)
        return [B.pack_tri(v) for v in V], exc, nel


def test_huckel_orbitals_orthonormal_and_occupied_space(oracle):
    bas = small_system("cc-pvdz", nw=2)
    S, T, V = int1e(bas)
    C, eps = huckel_guess(bas, 10)
    assert C.shape == (bas.nbf, bas.nbf)
    assert np.abs(C.T @ S @ C - np.eye(bas.nbf)).max() < 1e-12
    D = 2 * C[:, :10] @ C[:, :10].T
    assert np.sum(D * S) == pytest.approx(20.0, abs=1e-10)
    # the occupied Hueckel levels: two O 1s cores near the table's -20.7 Eh, valence above
    assert np.all(eps[:2] < -20.0) and np.all(eps[2:10] > -3.0)


def test_huckel_minimal_basis_energies_match_table(oracle):
    """Diagonal of the Hueckel operator = the orbital-energy table of huckel_lut.F90
    (H 1s -0.5, O 1s -20.7, 2s -1.244, 2p -0.632); a single atom has no
    off-diagonal coupling within a shell."""
    Z, xyz = np.array([8]), np.zeros((1, 3))
    mini = B.build_basis("mini_huckel", Z, xyz)
    with open(os.path.join(ROOT, "openqp_amd", "data", "huckel.json")) as fh:
        params = json.load(fh)
    _, e = huckel_minimal(mini, params)
    assert sorted(np.round(e, 6))[-3:] == [pytest.approx(-0.632)] * 3


def test_huckel_guess_energy_below_core_guess(oracle):
    bas = small_system("6-31g(d)")
    S, T, V = int1e(bas)
    H = T + V
    enuc = B.nuclear_repulsion(bas.zn, bas.xyz)

    def energy(D):
        g = oracle.fock_jk(bas, B.pack_tri(D))[0]
        F = H + B.unpack_tri(g, bas.nbf)
        return 0.5 * (np.sum(D * F) + np.sum(D * H)) + enuc
    C, _ = huckel_guess(bas, 5)
    e_h = energy(2 * C[:, :5] @ C[:, :5].T)
    e_vals, U = np.linalg.eigh(S)
    X = U / np.sqrt(e_vals)
    _, Co = np.linalg.eigh(X.T @ H @ X)
    Cc = X @ Co
    e_c = energy(2 * Cc[:, :5] @ Cc[:, :5].T)
    assert e_h < e_c and e_h - (-76.01074651325918) < 0.5


def test_scf_kat_with_huckel_guess_and_diis_no_warnings(oracle):
    """OpenQP KAT examples/HF/H2O_RHF-HF_ENERGY.json (-76.01074651325918 Eh,
    6-31G*, Cartesian d) through the host loop, with every warning an error
    (the DIIS solve must never be ill-conditioned)."""
    from openqp_amd.scf import RHF
    bas = small_system("6-31g(d)")
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        scf = RHF(bas, _CPUBuilder(oracle, bas), nocc=5, conv=1e-7, econv=1e-10)
        r_h = scf.run()
        r_c = RHF(bas, _CPUBuilder(oracle, bas), nocc=5, conv=1e-7, econv=1e-10, guess="core").run()
    assert r_h.converged and r_c.converged
    assert r_h.energy == pytest.approx(-76.01074651325918, abs=1e-8)
    assert abs(r_h.energy - r_c.energy) < 1e-9
    assert r_h.iterations <= r_c.iterations


def test_diis_drops_dependent_vectors():
    from openqp_amd.scf import RHF
    rng = np.random.default_rng(0)
    e1 = rng.standard_normal((4, 4))
    errs = [e1, 2.0 * e1, e1 * (1 + 1e-15), 0.5 * rng.standard_normal((4, 4))]
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        c = RHF.diis_coefficients(errs)
    assert np.isfinite(c).all() and c.sum() == pytest.approx(1.0)


def test_kohn_sham_b3lyp5_through_the_host_loop(oracle):
    """Closed-shell Kohn-Sham in the host loop (Hueckel guess, incremental
    builds with exchange scale 0.2, DIIS) against the reference's
    examples/TDDFT/H2O_B3LYP5-TDDFT_ENERGY.json, -76.37092191486505 Eh, to the
    grid error (5e-6; measured 8e-7)."""
    from openqp_amd.scf import RHF
    bas = small_system("6-31g(d)")
    r = RHF(bas, _CPUBuilder(oracle, bas), nocc=5, conv=1e-7, econv=1e-10, functional="b3lyp5").run()
    assert r.converged
    assert abs(r.nelec_grid - 10.0) < 1e-5
    assert abs(r.energy - (-76.37092191486505)) < 5e-6


def test_uks_triplet_through_the_host_loop(oracle):
    """UKS triplet B3LYP (libxc's B3LYP, VWN RPA fit) / cc-pVDZ H2O of
    examples/TRAH/H2O_UHF-DFT_GRAD_TRAH.json (-76.1477466020111 Eh) through the
    unrestricted host loop, from the closed-shell Kohn-Sham orbitals (from the
    core or Hueckel guess DIIS settles on an excited triplet), to the grid
    error (5e-6; measured 5e-7); N_elec on the grid 10."""
    from openqp_amd.scf import RHF, UHF
    Z = np.array([8, 1, 1], dtype=np.int32)
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, -1.430522676341708, 1.109269235155327],
                    [0.0, 1.430522676341708, 1.109269235155327]])
    bas = B.build_basis("cc-pvdz", Z, xyz)
    r0 = RHF(bas, _CPUBuilder(oracle, bas), nocc=5, conv=1e-5, functional="b3lypv1r").run()
    r = UHF(bas, _CPUBuilder(oracle, bas), mult=3, conv=1e-7, econv=1e-10, functional="b3lypv1r").run(r0.orbitals)
    assert r.converged
    assert abs(r.nelec_grid - 10.0) < 1e-5
    assert abs(r.energy - (-76.1477466020111)) < 5e-6, r.energy
