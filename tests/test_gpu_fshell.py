"""Classes with f shells (cc-pVTZ, pure and Cartesian) on the device: the
run-time-class kernels for the Schwarz bounds, the J/K build (RHF and UHF)
and the 2e gradient, against the CPU oracle on the same inputs.

Bars: Schwarz bounds Q^2 = max |(ab|ab)| to 1e-9 relative or 1e-15
absolute; J/K |f - f_oracle| <= 1e-10 Eh
elementwise with the oracle's computed-quartet count; gradient
|g - g_oracle| <= 1e-8 Eh/bohr and translational invariance."""
import sys

import numpy as np
import pytest

from helpers import small_system, density_for

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _basis(nw, harmonic=None, seed=0):
    from openqp_amd import basis as B, molecules as M
    if nw == 1:
        Z, xyz = M.h2o_kat()
    else:
        Z, xyz = M.water_grid(nw, 1, 1, seed=seed)
    return B.build_basis("cc-pvtz", Z, xyz, harmonic=harmonic)


@pytest.mark.parametrize("nw,harm", [(1, None), (1, 0), (2, None)])
def test_fshell_jk_parity(oracle, nw, harm):
    from openqp_amd import lib as L
    _say("f-shell J/K parity", nw, "pure" if harm is None else "cartesian")
    bas = _basis(nw, harm)
    D, dp = density_for(oracle, bas, 5 * nw, seed=1)
    Qo = oracle.schwarz(bas)
    fo, st = oracle.fock_jk(bas, dp, Q=Qo)
    fb = L.FockBuilder(bas)
    Q = fb.schwarz()
    # Q^2 = max |(ab|ab)|: the integral to 1e-9 relative or 1e-15 absolute.
    # The oracle's Obara-Saika recursion and the device's Rys quadrature
    # differ by up to 2e-10 relative on inter-molecular (fs|fs) diagonals of
    # (H2O)2 (5.9e-4, T large); the screening decisions, which are all the
    # bounds feed, agree exactly (quartet counts below).
    dq = np.abs(Q ** 2 - Qo ** 2)
    ex = dq - 1e-9 * Qo ** 2
    i, j = np.unravel_index(np.argmax(ex), ex.shape)
    _say("  Schwarz max |dQ^2| excess", ex.max(), "at shells", i, j, "L", bas.am[i], bas.am[j],
         "Q^2", Qo[i, j] ** 2, Q[i, j] ** 2)
    f = fb.fock_jk(dp)
    err = np.abs(f - fo).max()
    _say("  max |f - f_oracle| =", err, "quartets", fb.last_stats.quartets_computed, "oracle", int(st[1]))
    assert np.all(dq <= 1e-9 * Qo ** 2 + 1e-15)
    assert err <= TOL
    assert fb.last_stats.quartets_computed == int(st[1])
    fb.close()


def test_fshell_uhf_parity(oracle):
    from openqp_amd import lib as L, basis as B
    _say("f-shell UHF parity")
    bas = _basis(1)
    S = oracle.int1e(bas)[0]
    ds = []
    for k, sd in ((5, 21), (4, 22)):
        rng = np.random.default_rng(sd)
        C = rng.standard_normal((bas.nbf, k))
        Lc = np.linalg.cholesky(C.T @ S @ C)
        C = C @ np.linalg.inv(Lc).T
        ds.append(B.pack_tri(C @ C.T))
    Q = oracle.schwarz(bas)
    ra, rb, st = oracle.fock_jk_uhf(bas, ds[0], ds[1], se=1.0, sc=1.0, cutoff=5e-11, Q=Q)
    fb = L.FockBuilder(bas)
    fa, fbb = fb.fock_jk_uhf(ds[0], ds[1])
    err = max(np.abs(fa - ra).max(), np.abs(fbb - rb).max())
    _say("  max |f - f_oracle| =", err)
    assert err <= TOL
    fb.close()


@pytest.mark.parametrize("nw,harm", [(1, None), (1, 0), (2, None)])
def test_fshell_grad2_parity(oracle, nw, harm):
    from openqp_amd import lib as L
    _say("f-shell grad2 parity", nw, "pure" if harm is None else "cartesian")
    bas = _basis(nw, harm, seed=2)
    D, dp = density_for(oracle, bas, 5 * nw, seed=6)
    go, st = oracle.grad2(bas, dp)
    fb = L.FockBuilder(bas)
    g = fb.grad2(dp)
    err = np.abs(g - go).max()
    _say("  max |g - g_oracle| =", err, "quartets", fb.last_stats.quartets_computed, "oracle", st)
    assert err <= 1e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    fb.close()


def test_fshell_mixed_precision_energy(oracle):
    """Mixed precision leaves every f class in FP64; the FP32 share of the
    other classes keeps E2 within 1 uEh of the FP64 build."""
    from openqp_amd import lib as L
    bas = _basis(2)
    D, dp = density_for(oracle, bas, 10, seed=3)
    fb = L.FockBuilder(bas)
    f64 = fb.fock_jk(dp)
    fb.set_precision(1, 1e-4)
    fmx = fb.fock_jk(dp)
    n32 = fb.last_stats.quartets_fp32
    idx = np.arange(1, bas.nbf + 1)
    dg = idx * (idx + 1) // 2 - 1

    def e2(f):
        return 0.5 * (2 * np.dot(dp, f) - np.dot(dp[dg], f[dg]))
    dE = abs(e2(fmx) - e2(f64))
    _say("  fp32 quartets", n32, "dE2", dE)
    assert n32 > 0
    assert dE <= 1e-6
    fb.close()


@pytest.mark.parametrize("harm", [None, 0])
def test_fshell_runtime_class_kernels_only(oracle, monkeypatch, harm):
    """OQPH_F_TEMPLATES=0 sends every f class to the run-time-class kernels
    (k_jkh, k_gradh); J/K and gradient keep the same bars."""
    from openqp_amd import lib as L
    monkeypatch.setenv("OQPH_F_TEMPLATES", "0")
    _say("f-shell run-time-class path", "pure" if harm is None else "cartesian")
    bas = _basis(1, harm)
    D, dp = density_for(oracle, bas, 5, seed=4)
    Qo = oracle.schwarz(bas)
    fo, st = oracle.fock_jk(bas, dp, Q=Qo)
    go, _ = oracle.grad2(bas, dp, Q=Qo)
    fb = L.FockBuilder(bas)
    f = fb.fock_jk(dp)
    err = np.abs(f - fo).max()
    g = fb.grad2(dp)
    gerr = np.abs(g - go).max()
    _say("  max |f - f_oracle| =", err, "max |g - g_oracle| =", gerr)
    assert err <= TOL
    assert fb.last_stats.quartets_computed > 0
    assert gerr <= 1e-8
    fb.close()


def test_fshell_multi_class_launch(oracle, monkeypatch):
    """The run-time classes of a build merged into one persistent launch per
    LDS tier (k_jkh_multi, the default) against one launch per class
    (OQPH_HL_MULTI=0) and the oracle: same G to 1e-12, same quartet counts,
    and with every f class on the run-time-class kernel the merged launch
    covers several classes."""
    from openqp_amd import lib as L
    monkeypatch.setenv("OQPH_F_TEMPLATES", "0")
    bas = _basis(2, None)
    D, dp = density_for(oracle, bas, 10, seed=14)
    fo, st = oracle.fock_jk(bas, dp)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("OQPH_HL_MULTI", mode)
        fb = L.FockBuilder(bas)
        f = fb.fock_jk(dp)
        prof = fb.last_profile()
        out[mode] = (f, fb.last_stats.quartets_computed, fb.last_stats.quartets_screened,
                     [p["cls"] for p in prof if p["kernel"] == "k_jkh_multi"])
        fb.close()
    for mode in ("0", "1"):
        err = np.abs(out[mode][0] - fo).max()
        _say("  OQPH_HL_MULTI", mode, "max |f - f_oracle| =", err, "multi launches", out[mode][3])
        assert err <= TOL
        assert out[mode][1] == int(st[1])
    assert np.abs(out["0"][0] - out["1"][0]).max() <= 1e-12
    assert out["0"][1:3] == out["1"][1:3]
    assert out["0"][3] == []
    assert out["1"][3] and max(c - 40000 for c in out["1"][3]) >= 2


@pytest.mark.parametrize("name,nw", [("cc-pvdz", 2), ("cc-pvtz", 1)])
def test_routec_seam_symbols(oracle, name, nw):
    """Both Route-C symbols called by reference as OpenQP's seam calls them
    (routec_bridge.F90:33-61), on the default session: J/K against the oracle,
    the 2e gradient at the session geometry and, through the xyz argument, at a
    displaced geometry (the library refreshes its geometry)."""
    from openqp_amd import lib as L
    from openqp_amd import basis as B
    if name == "cc-pvtz":
        bas = _basis(nw)
    else:
        bas = small_system(name, nw=nw)
    D, dp = density_for(oracle, bas, 5 * nw, seed=7)
    fb = L.FockBuilder(bas)
    fb.make_default()
    f, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 0
    fo, _ = oracle.fock_jk(bas, dp, Q=oracle.schwarz(bas))
    assert np.abs(f - fo).max() <= TOL
    g, info = L.routec_grad2(dp, bas.xyz, bas.nbf)
    assert info == 0
    go, _ = oracle.grad2(bas, dp)
    _say("  seam", name, "max |f - f_oracle|", np.abs(f - fo).max(), "max |g - g_oracle|", np.abs(g - go).max())
    assert np.abs(g - go).max() <= 1e-8
    xyz2 = bas.xyz.copy()
    xyz2[0, 1] += 0.05
    g2, info = L.routec_grad2(dp, xyz2, bas.nbf)
    assert info == 0
    go2, _ = oracle.grad2(bas.with_coords(xyz2), dp)
    _say("  displaced geometry max |g - g_oracle|", np.abs(g2 - go2).max())
    assert np.abs(g2 - go2).max() <= 1e-8
    assert np.abs(g2 - g).max() > 1e-6   # the displacement is seen
    _, info = L.routec_grad2(dp, xyz2, bas.nbf + 1)   # mismatched nbf declines
    assert info != 0
    fb.close()
