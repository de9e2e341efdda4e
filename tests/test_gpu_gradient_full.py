"""Full RHF nuclear gradient: one-electron terms (host code in liboqph),
two-electron adjoint (device) and nuclear repulsion, against OpenQP's own
known answer (examples/HF/H2O_RHF-HF_GRADIENT.json, H2O RHF/6-31G*) and
against central finite differences of the SCF energy (1e-4 bohr)."""
import json
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, small_system

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_grad1e_matches_finite_differences():
    """sum D dH/dR - sum W dS/dR against central differences of
    sum D H(R) - sum W S(R) for random symmetric D, W (Cartesian and pure d)."""
    from openqp_amd import lib as L
    for name in ("6-31g(d)", "cc-pvdz"):
        bas = small_system(name, nw=2, seed=3)
        n = bas.nbf
        rng = np.random.default_rng(0)
        A = rng.standard_normal((n, n))
        D = (A + A.T) * 0.1
        A = rng.standard_normal((n, n))
        W = (A + A.T) * 0.1
        g = L.grad1e(bas, D, W)

        def energy(xyz):
            S, T, V = L.int1e(bas.with_coords(xyz))
            return np.sum(D * (T + V)) - np.sum(W * S)

        h = 1e-4
        err = 0.0
        for a in range(bas.natm):
            for d in range(3):
                xp = bas.xyz.copy(); xp[a, d] += h
                xm = bas.xyz.copy(); xm[a, d] -= h
                err = max(err, abs((energy(xp) - energy(xm)) / (2 * h) - g[a, d]))
        _say(name, "max |grad1e - fd| =", err)
        assert err < 1e-7


def _scf(bas, conv=1e-9):
    from openqp_amd.lib import FockBuilder
    from openqp_amd.scf import RHF
    fb = FockBuilder(bas, cutoff=1e-14, grad_cutoff=1e-14)
    res = RHF(bas, fb, conv=conv, econv=1e-12, max_iter=200).run()
    assert res.converged
    return fb, res


def test_rhf_gradient_known_answer():
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.gradient import rhf_gradient
    kat = json.load(open(os.path.join(ROOT, "tests", "golden", "h2o_rhf_gradient_kat.json")))
    Z, xyz = M.h2o_kat()
    assert np.abs(xyz.reshape(-1) - np.array(kat["coord_bohr"])).max() < 1e-6
    bas = B.build_basis("6-31g(d)", Z, xyz)
    fb, res = _scf(bas)
    g = rhf_gradient(bas, res.orbitals, res.orbital_energies, 5, fb)["total"]
    ref = np.array(kat["grad"]).reshape(3, 3)
    _say("E", res.energy, "ref", kat["energy"])
    _say("grad", g.reshape(-1))
    _say("ref ", ref.reshape(-1))
    assert abs(res.energy - kat["energy"]) < 1e-8
    assert np.abs(g - ref).max() < 1e-6
    fb.close()


def test_rhf_gradient_finite_differences():
    """Distorted water dimer, cc-pVDZ (pure d): analytic total gradient vs
    central differences of the converged SCF energy on six coordinates."""
    from openqp_amd.gradient import rhf_gradient
    bas0 = small_system("cc-pvdz", nw=2, seed=5)
    rng = np.random.default_rng(7)
    xyz0 = bas0.xyz + rng.normal(0.0, 0.05, bas0.xyz.shape)
    bas = bas0.with_coords(xyz0)
    nocc = int(round(bas.zn.sum())) // 2
    fb, res = _scf(bas)
    g = rhf_gradient(bas, res.orbitals, res.orbital_energies, nocc, fb)["total"]
    fb.close()
    h = 1e-4
    for a, d in [(0, 0), (1, 2), (2, 1), (3, 0), (4, 2), (5, 1)]:
        e = []
        for s in (+1, -1):
            x = xyz0.copy()
            x[a, d] += s * h
            fbs, rs = _scf(bas.with_coords(x))
            e.append(rs.energy)
            fbs.close()
        fd = (e[0] - e[1]) / (2 * h)
        _say("  coord", a, d, "analytic", g[a, d], "fd", fd)
        assert abs(g[a, d] - fd) < 1e-6
