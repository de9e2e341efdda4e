"""The analytic-gradient workload, C20H42 / 6-31G(d) (linear alkane, seeded
+-0.05 A noise): the device 2e gradient against the oracle on a small alkane
and against the oracle's fixture on all 186 coordinates of C20H42; the total
RHF gradient against central finite differences (1e-4 bohr) of the FP64 SCF
energy on 6 seeded coordinates; the gradient at the mixed-precision SCF
solution against the FP64 one on every coordinate."""
import os
import sys
import time

import numpy as np
import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _alkane_basis(nc, seed=0):
    from openqp_amd import basis as B, molecules as M
    Z, xyz = M.alkane(nc, seed=seed)
    return B.build_basis("6-31g(d)", Z, xyz)


def test_grad2_small_alkane_matches_oracle(oracle):
    """C4H10 / 6-31G(d): device 2e gradient at a seeded density <= 1e-8 against the
    oracle's grd2 restatement, every coordinate."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.lib import FockBuilder, int1e
    bas = _alkane_basis(4, seed=3)
    S, _, _ = int1e(bas)
    dp = B.pack_tri(M.random_density(S, int(round(bas.zn.sum())) // 2, seed=4))
    fb = FockBuilder(bas)
    g = fb.grad2(dp)
    go, _ = oracle.grad2(bas, dp)
    err = np.abs(g - go).max()
    _say("  C4H10 max |g - g_oracle| =", err)
    assert err <= 1e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-9
    fb.close()


def test_grad2_c20h42_fixture():
    """C20H42 / 6-31G(d), all 62 atoms x 3: device 2e gradient at the seeded density
    against the oracle's values in tests/golden/c20h42_631gd_grad2.npz (<= 1e-8)."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.lib import FockBuilder, int1e
    fx = np.load(os.path.join(ROOT, "tests", "golden", "c20h42_631gd_grad2.npz"))
    bld, bname, nocc = M.WORKLOADS["c20h42_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    assert np.abs(bas.xyz - fx["xyz"]).max() < 1e-12
    S, _, _ = int1e(bas)
    dp = B.pack_tri(M.random_density(S, int(fx["nocc"]), seed=int(fx["seed"])))
    assert abs(float(np.sum(dp * dp)) - float(fx["d_checksum"])) < 1e-9 * float(fx["d_checksum"])
    fb = FockBuilder(bas)
    g = fb.grad2(dp)
    err = np.abs(g - fx["de"]).max()
    _say("  C20H42 max |g - g_fixture| =", err, "over", g.size, "coordinates; quartets",
         fb.last_stats.quartets_computed)
    assert g.size == 186
    assert err <= 1e-8
    fb.close()


def _scf(bas, precision=0, fp32_cutoff=0.0, D0=None, cutoff=1e-13):
    from openqp_amd.lib import FockBuilder
    from openqp_amd.scf import RHF
    fb = FockBuilder(bas, cutoff=cutoff, grad_cutoff=cutoff, precision=precision, fp32_cutoff=fp32_cutoff)
    res = RHF(bas, fb, conv=1e-8, econv=1e-11, max_iter=200).run(D0)
    assert res.converged
    return fb, res


@pytest.fixture(scope="module")
def c20():
    from openqp_amd.gradient import rhf_gradient
    from openqp_amd import molecules as M, basis as B
    bld, bname, nocc = M.WORKLOADS["c20h42_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    t = time.perf_counter()
    fb, res = _scf(bas)
    g = rhf_gradient(bas, res.orbitals, res.orbital_energies, nocc, fb)["total"]
    fb.close()
    _say("  C20H42 FP64 SCF", res.energy, res.iterations, "iterations", f"{time.perf_counter() - t:.1f} s")
    return bas, nocc, res, g


def test_c20h42_gradient_finite_differences(c20):
    """Total RHF gradient (1e host terms + 2e device adjoint + nuclear) against
    central differences (h = 1e-4 bohr) of the converged FP64 SCF energy on six
    seeded coordinates: <= 1e-6 Eh/bohr."""
    bas, nocc, res, g = c20
    rng = np.random.default_rng(20)
    coords = [(int(a), int(d)) for a, d in zip(rng.choice(bas.natm, 6, replace=False), rng.integers(0, 3, 6))]
    h = 1e-4
    for a, d in coords:
        e = []
        for s in (+1, -1):
            x = bas.xyz.copy()
            x[a, d] += s * h
            fbs, rs = _scf(bas.with_coords(x), D0=res.density)
            e.append(rs.energy)
            fbs.close()
        fd = (e[0] - e[1]) / (2 * h)
        _say("  atom", a, "axis", d, "analytic", g[a, d], "fd", fd, "diff", abs(g[a, d] - fd))
        assert abs(g[a, d] - fd) < 1e-6


def test_c20h42_mixed_precision_gradient(c20):
    """SCF with mixed-precision builds (FP32 below a bound of 1e-4, FP64
    accumulation): |dE_SCF| <= 1 uEh and max |dgrad| <= 1e-6 Eh/bohr against the
    FP64 solution, on all 186 coordinates."""
    from openqp_amd.gradient import rhf_gradient
    bas, nocc, res, g = c20
    fb, rm = _scf(bas, precision=1, fp32_cutoff=1e-4)
    n32 = fb.last_stats.quartets_fp32
    fb.set_precision(0)
    gm = rhf_gradient(bas, rm.orbitals, rm.orbital_energies, nocc, fb)["total"]
    fb.close()
    dE, dg = abs(rm.energy - res.energy), np.abs(gm - g).max()
    _say("  mixed SCF: |dE| =", dE, "max |dgrad| =", dg, "fp32 quartets in the last build", n32)
    assert dE <= 1e-6
    assert dg <= 1e-6
