"""Device parity of the range-separated (CAM) build, of OpenQP's MPI rank
shares (one rank's quartets on one GPU, also at the full benchmark sizes),
of the mixed-precision path on a translated cluster, of the seam after a
geometry change, and of the fault-injection declines.  Through the C ABI;
|f_gpu - f_oracle| <= 1e-10 elementwise."""
import os
import sys
import time

import numpy as np
import pytest

from helpers import small_system, workload, density_for

pytestmark = pytest.mark.gpu
TOL = 1e-10
CAM_B3LYP = dict(alpha=0.19, beta=0.46, mu=0.33)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    from openqp_amd import lib
    return lib


@pytest.mark.parametrize("name,nw,nocc", [("cc-pvdz", 2, 10), ("6-31g(d)", 3, 15), ("cc-pvtz", 1, 5)])
def test_cam_parity(L, oracle, name, nw, nocc):
    """CAM-B3LYP parameters: 1/r pass (Coulomb 1, exchange 0.19) plus the
    erf(0.33 r)/r pass (exchange 0.46), each screened with its own bounds."""
    _say("CAM parity", name, nw)
    bas = small_system(name, nw=nw, seed=2)
    D, dp = density_for(oracle, bas, nocc, seed=21)
    fo, so = oracle.fock_jk_cam(bas, dp, **CAM_B3LYP)
    fb = L.FockBuilder(bas)
    f = fb.fock_jk_cam(dp, **CAM_B3LYP)
    err = np.abs(f - fo).max()
    _say("  max |f - f_oracle| =", err, "quartets", fb.last_stats.quartets_computed, int(so[1]))
    assert err <= TOL
    assert fb.last_stats.quartets_computed == int(so[1])
    # the attenuated pass alone, with a Coulomb part: beta_coulomb = 1
    fo2, _ = oracle.fock_jk_cam(bas, dp, alpha=0.0, beta=1.0, mu=0.7, alpha_coulomb=0.0, beta_coulomb=1.0)
    f2 = fb.fock_jk_cam(dp, alpha=0.0, beta=1.0, mu=0.7, alpha_coulomb=0.0, beta_coulomb=1.0)
    assert np.abs(f2 - fo2).max() <= TOL
    fb.close()


def test_cam_uhf_parity(L, oracle):
    from openqp_amd import basis as B
    bas = small_system("cc-pvdz", nw=2, seed=4)
    S = oracle.int1e(bas)[0]
    ds = []
    for k, sd in ((10, 31), (9, 32)):
        rng = np.random.default_rng(sd)
        C = rng.standard_normal((bas.nbf, k))
        Lc = np.linalg.cholesky(C.T @ S @ C)
        C = C @ np.linalg.inv(Lc).T
        ds.append(B.pack_tri(C @ C.T))
    d = np.concatenate(ds)
    fo, _ = oracle.fock_jk_cam(bas, d, nfocks=2, **CAM_B3LYP)
    fb = L.FockBuilder(bas)
    f = fb.fock_jk_cam(d, nfocks=2, **CAM_B3LYP)
    err = np.abs(f - fo).max()
    _say("  CAM UHF max |f - f_oracle| =", err)
    assert err <= TOL
    fb.close()


@pytest.mark.parametrize("size", [2, 3, 5])
def test_rank_shares_small(L, oracle, size):
    """Each rank's share equals the oracle's share of OpenQP's cost-sorted pair
    map (same bounds, so the same map), with the same quartet count; the
    shares sum to the full build."""
    bas = small_system("cc-pvdz", nw=4, seed=5)
    D, dp = density_for(oracle, bas, 20, seed=6)
    fb = L.FockBuilder(bas)
    Q = fb.schwarz()
    full = fb.fock_jk(dp)
    nfull = fb.last_stats.quartets_computed
    tot = np.zeros_like(full)
    nq = 0
    for r in range(size):
        fb.set_share(r, size)
        f = fb.fock_jk(dp)
        fo, st = oracle.fock_jk(bas, dp, Q=Q, rank=r, size=size)
        err = np.abs(f - fo).max()
        _say("  share", r, "of", size, "max |f - f_oracle| =", err, "quartets", fb.last_stats.quartets_computed,
             int(st[1]))
        assert err <= TOL
        assert fb.last_stats.quartets_computed == int(st[1])
        tot += f
        nq += fb.last_stats.quartets_computed
    fb.set_share(0, 1)
    assert np.abs(tot - full).max() <= 1e-12
    assert nq == nfull
    fb.close()


@pytest.mark.parametrize("wl,size", [("h2o64_ccpvdz", 32), ("h2o32_631gd", 8), ("h2o16_ccpvtz", 16)])
def test_rank_share_full_size(L, oracle, wl, size):
    """One rank's share of the full benchmark workloads on the device against
    the oracle's share on the host cores (a ~1/size sample of the build):
    elementwise parity at the sizes the benchmark measures."""
    bas, nocc = workload(wl)
    from openqp_amd import lib as LL
    S, _, _ = LL.int1e(bas)
    from openqp_amd import molecules as M, basis as B
    dp = B.pack_tri(M.random_density(S, nocc, seed=0))
    fb = L.FockBuilder(bas)
    Q = fb.schwarz()
    fb.set_share(1, size)
    f = fb.fock_jk(dp)
    nq = fb.last_stats.quartets_computed
    _say("  share 1 of", size, wl, "device quartets", nq)
    t = time.perf_counter()
    fo, st = oracle.fock_jk(bas, dp, Q=Q, rank=1, size=size)
    err = np.abs(f - fo).max()
    _say("  oracle share", f"{time.perf_counter() - t:.1f} s", "max |f - f_oracle| =", err, "quartets", int(st[1]))
    assert err <= TOL
    assert nq == int(st[1])
    fb.close()


def test_mixed_precision_translation_invariant(L, oracle):
    """FP32 quadrature sees only differences formed in FP64: the mixed build of a
    cluster moved by 50 bohr matches the unmoved one, and both stay within
    1 uEh of the FP64 oracle."""
    from openqp_amd.autograd import packed_trace
    bas = small_system("cc-pvdz", nw=4, seed=9)
    D, dp = density_for(oracle, bas, 20, seed=10)
    fo, _ = oracle.fock_jk(bas, dp)
    e0 = packed_trace(dp, fo, bas.nbf)
    out = []
    for shift in (0.0, 50.0):
        b2 = bas.with_coords(bas.xyz + np.array([shift, -0.6 * shift, 0.8 * shift]))
        fb = L.FockBuilder(b2, precision=1, fp32_cutoff=1e-4)
        f = fb.fock_jk(dp)
        out.append(packed_trace(dp, f, bas.nbf))
        _say("  shift", shift, "dE2 vs FP64", abs(out[-1] - e0), "fp32 quartets", fb.last_stats.quartets_fp32)
        assert abs(out[-1] - e0) <= 1e-6
        fb.close()
    assert abs(out[1] - out[0]) <= 1e-9


def test_seam_after_geometry_change(L, oracle):
    """The default session used by routec_fock_jk follows oqph_session_set_geometry:
    after moving the atoms, the seam builds at the new geometry (pair data and
    Schwarz bounds rebuilt)."""
    bas = small_system("cc-pvdz", nw=2)
    D, dp = density_for(oracle, bas, 10, seed=8)
    fb = L.FockBuilder(bas)
    fb.make_default()
    f1, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 0
    xyz2 = bas.xyz.copy()
    xyz2[1] += np.array([0.2, -0.1, 0.05])
    fb.set_geometry(xyz2)
    f2, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 0
    fo2, _ = oracle.fock_jk(bas.with_coords(xyz2), dp)
    _say("  seam after move: max |f - f_oracle(B)| =", np.abs(f2 - fo2).max())
    assert np.abs(f2 - fo2).max() <= TOL
    assert np.abs(f2 - f1).max() > 1e-4
    fb.close()


def test_fault_injection_declines(L, oracle, monkeypatch):
    """OQPH_FAULT makes every build decline (info != 0) so the caller's native path
    runs; nothing aborts."""
    bas = small_system("cc-pvdz", nw=1)
    D, dp = density_for(oracle, bas, 5, seed=1)
    fb = L.FockBuilder(bas)
    fb.make_default()
    monkeypatch.setenv("OQPH_FAULT", "decline")
    _, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 1
    with pytest.raises(L.OqphError):
        fb.fock_jk(dp)
    monkeypatch.setenv("OQPH_FAULT", "oom")
    _, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 3
    monkeypatch.delenv("OQPH_FAULT")
    f, info = L.routec_fock_jk(dp, bas.nbf)
    assert info == 0
    fb.close()


def test_kernel_resources_and_budget(L):
    """Every class-templated kernel's compiled resources are readable without a
    launch; the sessions route any whose private segment exceeds the 8704-byte
    budget to the run-time-class kernels (the (dd|dd) wavefront kernel does)."""
    classes = [(a, b, c, d) for a in range(4) for b in range(a + 1) for c in range(a + 1) for d in range(c + 1)
               if (a * 10 + b) >= (c * 10 + d)]
    seen = 0
    over = []
    for kind in (0, 1, 2, 3):
        for (a, b, c, d) in classes:
            if kind == 3 and (c, d) != (a, b):
                continue
            for pure in (False, True):
                r = L.kernel_resources(kind, a, b, c, d, pure)
                if r is None:
                    continue
                seen += 1
                if r[0] > 8704:
                    over.append((kind, a, b, c, d, pure, r[0]))
    _say("  templated kernels", seen, "over the private budget:", over)
    assert seen > 100
    # the budget check must see the known heavy template
    assert any(o[:5] == (0, 2, 2, 2, 2) for o in over)
