"""MRSF response integral pass on the device (oqph_mrsf_int2) against the
oracle's restatement of int2_mrsf_data_t (tdhf_mrsf_lib.F90:58-308)."""
import sys

import numpy as np
import pytest

from helpers import small_system

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _d3(bas, nf, seed):
    """Non-symmetric transition densities d3 (C order [nu, mu, c, v]) of the
    size of MRSF's: C_o x C_v^T products of random orbitals."""
    rng = np.random.default_rng(seed)
    n = bas.nbf
    d3 = np.zeros((n, n, 7, nf))
    for v in range(nf):
        for c in range(7):
            a = rng.standard_normal((n, 3)) * 0.3
            b = rng.standard_normal((n, 3)) * 0.3
            d3[:, :, c, v] = a @ b.T
    return d3


@pytest.mark.parametrize("name,nw,pure", [("6-31g(d)", 2, None), ("cc-pvdz", 2, None), ("cc-pvtz", 1, None),
                                          ("cc-pvtz", 1, 0)])
def test_mrsf_int2_matches_oracle(oracle, name, nw, pure):
    from openqp_amd import lib as L
    _say("mrsf_int2", name, nw, "pure" if pure is None else "cartesian")
    from openqp_amd import basis as B, molecules as M
    if pure is None:
        bas = small_system(name, nw=nw, seed=2)
    else:
        Z, xyz = M.h2o_kat()
        bas = B.build_basis(name, Z, xyz, harmonic=pure)
    d3 = _d3(bas, 2, seed=5)
    fo, so = oracle.mrsf_int2(bas, d3, ncoul=4, se=0.8, sc=0.8)
    fb = L.FockBuilder(bas)
    f3 = fb.mrsf_int2(d3, ncoul=4, se=0.8, sc=0.8)
    st = fb.last_stats
    err = np.abs(f3 - fo).max()
    _say("  max |f3 - f3_oracle| =", err, " quartets", st.quartets_computed, "oracle", so[1])
    assert err <= 1e-10
    assert st.quartets_computed == so[1]
    # the Fock build of the same session still works after the response pass
    D = d3[:, :, 6, 0] + d3[:, :, 6, 0].T
    from openqp_amd import basis as B
    dp = B.pack_tri(D)
    g = fb.fock_jk(dp)
    go, _ = oracle.fock_jk(bas, dp)
    assert np.abs(g - go).max() <= 1e-10
    fb.close()


def test_mrsf_int2_symmetric_density_is_exchange_build(oracle):
    """A symmetric exchange-only component reproduces the RHF exchange build
    (se = 2, sc = 0 gives packed -K) on the device."""
    from openqp_amd import basis as B, lib as L
    bas = small_system("cc-pvdz", nw=2, seed=4)
    n = bas.nbf
    rng = np.random.default_rng(9)
    A = rng.standard_normal((n, 4)) * 0.4
    D = A @ A.T
    d3 = np.zeros((n, n, 7, 1))
    d3[:, :, 6, 0] = D
    fb = L.FockBuilder(bas)
    f3 = fb.mrsf_int2(d3, se=1.0, sc=1.0)
    fr, _ = oracle.fock_jk(bas, B.pack_tri(D), se=2.0, sc=0.0)
    assert np.abs(B.pack_tri(f3[:, :, 6, 0]) - fr).max() <= 1e-10
    fb.close()


@pytest.mark.parametrize("name,nw,kind", [("6-31g(d)", 1, 1), ("6-31g(d)", 1, 3), ("cc-pvdz", 2, 1), ("cc-pvdz", 2, 3)])
def test_mrsf_sigma_session_matches_oracle(oracle, name, nw, kind):
    """routec_sig_init / set_scale / iter / free through the C ABI, by
    reference as OpenQP's routec_sig module calls them, against the oracle's
    restatement of mrsfcbc -> int2 -> mrsfmntoia + mrsfesum; two iterations,
    the second with fewer vectors; 1e-10 relative to the largest element."""
    from oracle import mrsf_sigma as MS
    from openqp_amd import lib as L
    _say("MRSF sigma", name, nw, "kind", kind)
    bas = small_system(name, nw=nw, seed=3)
    n = bas.nbf
    S, _, _ = L.int1e(bas)
    rng = np.random.default_rng(7)
    w, U = np.linalg.eigh(S)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    va = U @ np.diag(w ** -0.5) @ U.T @ Q
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    vb = U @ np.diag(w ** -0.5) @ U.T @ Q2
    fa = rng.standard_normal((n, n)); fa = fa + fa.T
    fb = rng.standard_normal((n, n)); fb = fb + fb.T
    nocca = 5 * nw + 1
    noccb = nocca - 2
    fbld = L.FockBuilder(bas)
    fbld.make_default()
    sig = L.SigmaSession(va, vb, fa, fb, nocca, noccb, kind, scale=0.6)
    try:
        for nv in (3, 2):
            B = rng.standard_normal((nocca * (n - noccb), nv))
            got = sig.iter(B)
            ref = MS.sigma(bas, va, vb, fa, fb, nocca, noccb, kind, 0.6, B)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            _say("  nv", nv, "max rel |sigma - sigma_oracle| =", err)
            assert err <= 1e-10
    finally:
        sig.close()
        fbld.close()


def _grad_inputs(n, rng):
    A = rng.standard_normal((n, n)) * 0.3
    Bm = rng.standard_normal((n, n)) * 0.3
    d = np.array([(A @ A.T / n).T, (Bm @ Bm.T / n).T])
    Pa = rng.standard_normal((n, n)) * 0.05
    Pb = rng.standard_normal((n, n)) * 0.05
    p = np.array([(Pa + Pa.T).T, (Pb + Pb.T).T])
    spc = rng.standard_normal((n, n, 7)) * 0.1
    return d, p, spc


@pytest.mark.parametrize("name,nw,mrst", [("6-31g(d)", 1, 1), ("6-31g(d)", 1, 3), ("cc-pvdz", 2, 1)])
def test_mrsf_grad2_matches_oracle(oracle, name, nw, mrst):
    """MRSF 2e gradient (routec_grad2_mrsf, by reference through the C ABI,
    and oqph_grad2_mrsf) against the oracle's grd2_mrsf restatement, 1e-8;
    afterwards the ground-state gradient of the same session is unchanged."""
    import ctypes as C
    from openqp_amd import basis as B, lib as L
    _say("MRSF gradient", name, nw, "mrst", mrst)
    bas = small_system(name, nw=nw, seed=6)
    n = bas.nbf
    rng = np.random.default_rng(40 + mrst)
    d, p, spc = _grad_inputs(n, rng)
    hs, hs2, cs, qf = 0.7, 0.5, 1.0, (0.9, 1.1, 0.8)
    go, so = oracle.grad2_mrsf(bas, d, p, spc, hs, hs2, cs, qf, mrst)
    fb = L.FockBuilder(bas)
    fb.make_default()
    g = fb.grad2_mrsf(d, p, spc, hs, hs2, cs, qf, mrst)
    err = np.abs(g - go).max()
    _say("  max |g - g_oracle| =", err, "quartets", fb.last_stats.quartets_computed, "oracle", so[2])
    assert err <= 1e-8
    assert fb.last_stats.quartets_computed == so[2]
    # the reference's by-reference ABI
    lib = L.load()
    dp_ = C.POINTER(C.c_double)
    xyz = np.ascontiguousarray(np.asarray(bas.xyz, np.float64).reshape(-1))
    de = np.zeros(3 * bas.natm)
    arrs = [np.ascontiguousarray(x, np.float64) for x in (d, p, spc)]
    sc = np.array(qf, np.float64)
    info = C.c_int(-1)
    lib.routec_grad2_mrsf(*[a.ctypes.data_as(dp_) for a in arrs], xyz.ctypes.data_as(dp_), de.ctypes.data_as(dp_),
                          C.byref(C.c_int(n)), C.byref(C.c_int(bas.natm)), C.byref(C.c_double(hs)),
                          C.byref(C.c_double(hs2)), C.byref(C.c_double(cs)), sc.ctypes.data_as(dp_),
                          C.byref(C.c_int(mrst)), C.byref(info))
    assert info.value == 0
    assert np.abs(de.reshape(-1, 3) - go).max() <= 1e-8
    # ground-state gradient after the MRSF one: class templates again, same result as the oracle
    D = d[0].T + d[1].T
    dpk = B.pack_tri(0.5 * (D + D.T))
    g0, _ = oracle.grad2(bas, dpk)
    assert np.abs(fb.grad2(dpk) - g0).max() <= 1e-8
    fb.close()


def test_mrsf_seams_on_h2o8_631g(oracle):
    """The three MRSF paths on the (H2O)8/6-31G workload (104 bf, the
    CPU-runnable benchmark system) against the oracle: response pass
    (1e-10, same counts), sigma iteration (1e-10 relative) and MRSF 2e gradient
    (1e-8, same counts)."""
    from oracle import mrsf_sigma as MS
    from openqp_amd import basis as B, lib as L, molecules as M
    bld, bname, nocc = M.WORKLOADS["h2o8_631g"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    n = bas.nbf
    _say("MRSF seams on (H2O)8/6-31G, nbf", n)
    fb = L.FockBuilder(bas)
    fb.make_default()
    d3 = _d3(bas, 2, seed=11)
    fo, so = oracle.mrsf_int2(bas, d3, ncoul=4, se=0.5, sc=0.5)
    f3 = fb.mrsf_int2(d3, ncoul=4, se=0.5, sc=0.5)
    _say("  response pass max |df3| =", np.abs(f3 - fo).max())
    assert np.abs(f3 - fo).max() <= 1e-10
    assert fb.last_stats.quartets_computed == so[1]
    S, _, _ = L.int1e(bas)
    rng = np.random.default_rng(12)
    w, U = np.linalg.eigh(S)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    va = U @ np.diag(w ** -0.5) @ U.T @ Q
    fa = np.diag(np.sort(rng.standard_normal(n)))
    nocca, noccb = nocc + 1, nocc - 1
    sig = L.SigmaSession(va, va, fa, fa, nocca, noccb, 1, scale=0.5)
    bv = rng.standard_normal((nocca * (n - noccb), 2))
    got = sig.iter(bv)
    ref = MS.sigma(bas, va, va, fa, fa, nocca, noccb, 1, 0.5, bv)
    sig.close()
    _say("  sigma max rel =", np.abs(got - ref).max() / np.abs(ref).max())
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    d, p, spc = _grad_inputs(n, rng)
    go, sg = oracle.grad2_mrsf(bas, d, p, spc, 0.5, 0.5, 1.0, (1.0, 1.0, 1.0), 1)
    g = fb.grad2_mrsf(d, p, spc, 0.5, 0.5, 1.0, (1.0, 1.0, 1.0), 1)
    _say("  MRSF gradient max |dg| =", np.abs(g - go).max())
    assert np.abs(g - go).max() <= 1e-8
    assert fb.last_stats.quartets_computed == sg[2]
    fb.close()
