"""Device adjoint (2e gradient) and host SCF on the device Fock builder.

* grad2 on the device vs the oracle's restatement of hf_2e_grad/grd2:
  max |dg| <= 1e-8 Eh/bohr (FP64);
* the torch.autograd.Function backward equals the device adjoint and
  matches central differences (1e-4 bohr) of E2 = 0.5 Tr D G[D] built on the
  device;
* RHF/6-31G* H2O through the device Fock builder reproduces OpenQP's known
  answer E = -76.01074651325918 (examples/HF/H2O_RHF-HF_ENERGY.json) to 1e-8.
"""
import sys

import numpy as np
import pytest

from helpers import small_system, density_for

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,nw,nocc", [("6-31g(d)", 1, 5), ("cc-pvdz", 2, 10), ("6-31g(d)", 3, 15)])
def test_grad2_parity(oracle, name, nw, nocc):
    from openqp_amd import lib as L
    _say("grad2 parity", name, nw)
    bas = small_system(name, nw=nw, seed=2)
    D, dp = density_for(oracle, bas, nocc, seed=6)
    go, st = oracle.grad2(bas, dp)
    fb = L.FockBuilder(bas)
    g = fb.grad2(dp)
    err = np.abs(g - go).max()
    _say("  max |g - g_oracle| =", err, "quartets", fb.last_stats.quartets_computed, "oracle", st)
    assert err <= 1e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    fb.close()


def test_autograd_backward_and_fd(oracle):
    import torch
    from openqp_amd import lib as L
    from openqp_amd.autograd import TwoElectronEnergy
    bas = small_system("cc-pvdz", nw=2, seed=4)
    D, dp = density_for(oracle, bas, 10, seed=8)
    fb = L.FockBuilder(bas, cutoff=1e-14, grad_cutoff=1e-14)
    x = torch.tensor(bas.xyz.copy(), dtype=torch.float64, requires_grad=True)
    e = TwoElectronEnergy.apply(x, fb, dp)
    (g,) = torch.autograd.grad(e, x)
    fb.set_geometry(bas.xyz)
    assert np.abs(g.numpy() - fb.grad2(dp)).max() < 1e-12
    h = 1e-4
    for a, d in [(0, 2), (1, 0), (3, 1), (5, 2), (2, 1), (4, 0)]:
        xp = bas.xyz.copy(); xp[a, d] += h
        xm = bas.xyz.copy(); xm[a, d] -= h
        ep = TwoElectronEnergy.apply(torch.tensor(xp), fb, dp).item()
        em = TwoElectronEnergy.apply(torch.tensor(xm), fb, dp).item()
        fd = (ep - em) / (2 * h)
        _say("  fd", a, d, g[a, d].item(), fd)
        assert abs(g[a, d].item() - fd) < 1e-7
    fb.close()


def test_scf_kat_h2o_631gs():
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.scf import RHF
    Z, xyz = M.h2o_kat()
    bas = B.build_basis("6-31g(d)", Z, xyz)
    res = RHF(bas, conv=1e-7, econv=1e-10).run()
    _say("SCF H2O/6-31G*", res.energy, res.iterations, res.converged)
    assert res.converged
    assert abs(res.energy - (-76.01074651325918)) < 1e-8


@pytest.mark.parametrize("mode", ["0", "300"])
def test_grad2_template_and_runtime_classes(oracle, monkeypatch, mode):
    """The gradient classes above 300 Cartesian components ((dp|dp), (dd|pp)
    and up) run on the run-time-class kernel by default; OQPH_GRAD_HL_NG=0
    keeps them on their class templates.  Both against the oracle, 1e-8."""
    from openqp_amd import lib as L
    monkeypatch.setenv("OQPH_GRAD_HL_NG", mode)
    _say("grad2, OQPH_GRAD_HL_NG", mode)
    bas = small_system("cc-pvdz", nw=2, seed=5)
    D, dp = density_for(oracle, bas, 10, seed=8)
    go, st = oracle.grad2(bas, dp)
    fb = L.FockBuilder(bas)
    g = fb.grad2(dp)
    err = np.abs(g - go).max()
    _say("  max |g - g_oracle| =", err)
    assert err <= 1e-8
    fb.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cut,tol", [(1e-2, 1e-6), (1e30, 1e-5)])
def test_grad2_mixed_precision(oracle, cut, tol):
    """Mixed-precision gradient: the template quartets with Q_ij Q_kl below
    `cut` run in FP32 (1e30: all of them).  Against the FP64 oracle; the FP64
    gradient of the same session is unchanged after switching back."""
    from openqp_amd import lib as L
    _say("grad2 mixed precision, cut", cut)
    bas = small_system("6-31g(d)", nw=3, seed=11)
    D, dp = density_for(oracle, bas, 15, seed=4)
    go, st = oracle.grad2(bas, dp)
    fb = L.FockBuilder(bas)
    g64 = fb.grad2(dp)
    fb.set_grad_precision(1, cut)
    g = fb.grad2(dp)
    err = np.abs(g - go).max()
    _say("  max |g_mixed - g_oracle| =", err, " FP64:", np.abs(g64 - go).max())
    assert err <= tol
    assert np.abs(g - g64).max() > 0.0  # the FP32 form ran
    fb.set_grad_precision(0, 0.0)
    assert np.abs(fb.grad2(dp) - g64).max() <= 1e-12
    fb.close()
