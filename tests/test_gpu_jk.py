"""Device J/K parity against the CPU oracle (and its committed golden
fixtures): |f_gpu - f_oracle| <= 1e-10 Eh elementwise (FP64), the tolerance
stated for this path.  Runs through the C ABI (liboqph)."""
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, small_system, workload, density_for

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    from openqp_amd import lib
    return lib


def _golden(name):
    return np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))


@pytest.mark.parametrize("name", ["h2o_631gs", "h2o2_ccpvdz", "h2o8_631g", "h2o_ccpvtz"])
def test_golden_parity(L, name):
    _say("golden parity", name)
    bas = {"h2o8_631g": lambda: workload("h2o8_631g")[0], "h2o_631gs": lambda: small_system("6-31g(d)"),
           "h2o2_ccpvdz": lambda: small_system("cc-pvdz", nw=2),
           "h2o_ccpvtz": lambda: small_system("cc-pvtz")}[name]()
    g = _golden(name)
    fb = L.FockBuilder(bas)
    Q = fb.schwarz()
    # Q = sqrt(max|(ab|ab)|): agreement to rounding of the element maxima
    rel = np.abs(Q - g["Q"]) / np.maximum(np.abs(g["Q"]), 1e-300)
    assert rel.max() < 1e-10, rel.max()
    f = fb.fock_jk(g["d"])
    err = np.abs(f - g["f"]).max()
    _say("  max |f - f_oracle| =", err, "stats", fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
    assert err <= TOL
    fb.close()


def test_screen_counts_match_oracle(L, oracle):
    bas, nocc = workload("h2o8_631g")
    g = _golden("h2o8_631g")
    fb = L.FockBuilder(bas)
    fb.fock_jk(g["d"])
    st = fb.last_stats
    # oracle stats: [nschwz, computed quartets, stored integrals]
    assert st.quartets_computed == int(g["stats"][1])
    fb.close()


@pytest.mark.parametrize("name", ["h2o8_631g", "h2o2_ccpvdz"])
def test_bra_row_kernel_matches_wave_kernel(L, name):
    """The bra-row kernel (K rows staged in LDS) and the wavefront-task kernel
    give the same G and the same screening statistics (OQPH_JK_ROWS=2 forces
    the bra-row kernel wherever its rows fit in LDS, =3 the same without the
    LDS copy of the density rows, =0 disables it)."""
    _say("bra-row vs wavefront kernel", name)
    bas = workload("h2o8_631g")[0] if name == "h2o8_631g" else small_system("cc-pvdz", nw=2)
    g = _golden(name)
    out = {}
    old = os.environ.get("OQPH_JK_ROWS")
    try:
        for mode in ("0", "2", "3"):
            os.environ["OQPH_JK_ROWS"] = mode
            fb = L.FockBuilder(bas)
            fb.schwarz()
            out[mode] = (fb.fock_jk(g["d"]), fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
            fb.close()
    finally:
        if old is None:
            os.environ.pop("OQPH_JK_ROWS", None)
        else:
            os.environ["OQPH_JK_ROWS"] = old
    for mode in ("0", "2", "3"):
        err = np.abs(out[mode][0] - g["f"]).max()
        _say("  mode", mode, "max |f - f_oracle| =", err)
        assert err <= TOL
    for mode in ("2", "3"):
        assert np.abs(out["0"][0] - out[mode][0]).max() <= 1e-12
        assert out["0"][1:] == out[mode][1:]


@pytest.mark.parametrize("basis,nw,nocc", [("cc-pvdz", 3, 15), ("cc-pvtz", 1, 5)])
def test_grouped_kernel_modes(L, oracle, basis, nw, nocc):
    """The grouped bra-row kernel (pairs whose shells share a primitive set:
    cc-pVXZ 1s/2s) against the oracle and against the segmented kernels:
    OQPH_JKG=0 (off), 1 (default class pairs), 2 (every class pair with a
    grouped kernel).  Same G to 1e-12, same computed and screened counts."""
    _say("grouped kernel modes", basis, nw)
    bas = small_system(basis, nw=nw, seed=9)
    D, dp = density_for(oracle, bas, nocc, seed=19)
    fo, so = oracle.fock_jk(bas, dp)
    out, kinds = {}, {}
    old = os.environ.get("OQPH_JKG")
    try:
        for mode in ("0", "1", "2"):
            os.environ["OQPH_JKG"] = mode
            fb = L.FockBuilder(bas)
            f = fb.fock_jk(dp)
            out[mode] = (f, fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
            kinds[mode] = {p["kernel"] for p in fb.last_profile()}
            fb.close()
    finally:
        if old is None:
            os.environ.pop("OQPH_JKG", None)
        else:
            os.environ["OQPH_JKG"] = old
    for mode in ("0", "1", "2"):
        err = np.abs(out[mode][0] - fo).max()
        _say("  OQPH_JKG", mode, "max |f - f_oracle| =", err, "quartets", out[mode][1], int(so[1]))
        assert err <= TOL
        assert out[mode][1] == int(so[1])
    for mode in ("1", "2"):
        assert np.abs(out["0"][0] - out[mode][0]).max() <= 1e-12
        assert out["0"][1:] == out[mode][1:]
    assert "k_jkg" not in kinds["0"]
    assert "k_jkg" in kinds["2"]


@pytest.mark.parametrize("name,nw,nocc", [("cc-pvdz", 3, 15), ("6-31g(d)", 4, 20)])
def test_parity_fresh_density(L, oracle, name, nw, nocc):
    _say("fresh parity", name, nw)
    bas = small_system(name, nw=nw, seed=3)
    D, dp = density_for(oracle, bas, nocc, seed=11)
    fo, st = oracle.fock_jk(bas, dp)
    fb = L.FockBuilder(bas)
    f = fb.fock_jk(dp)
    err = np.abs(f - fo).max()
    _say("  max err", err)
    assert err <= TOL
    # scale factors (se, sc) as fock_jk passes them
    fo2, _ = oracle.fock_jk(bas, dp, se=0.25, sc=0.7)
    assert np.abs(fb.fock_jk(dp, se=0.25, sc=0.7) - fo2).max() <= TOL
    fb.close()


def test_routec_symbol_matches_session(L):
    bas = small_system("cc-pvdz", nw=2)
    g = _golden("h2o2_ccpvdz")
    fb = L.FockBuilder(bas)
    fb.make_default()
    f, info = L.routec_fock_jk(g["d"], bas.nbf)
    assert info == 0
    assert np.abs(f - g["f"]).max() <= TOL
    f2, info2 = L.routec_fock_jk(g["d"], bas.nbf + 1)   # mismatched nbf declines
    assert info2 != 0
    fb.close()


def test_zero_and_linearity(L, oracle):
    bas = small_system("cc-pvdz", nw=2)
    fb = L.FockBuilder(bas, cutoff=1e-20)
    n = bas.ntri
    assert np.all(fb.fock_jk(np.zeros(n)) == 0.0)
    D, dp = density_for(oracle, bas, 10, seed=5)
    f1 = fb.fock_jk(dp)
    f2 = fb.fock_jk(2.5 * dp)
    assert np.abs(f2 - 2.5 * f1).max() < 1e-11
    fb.close()


def test_incremental_equals_full(L, oracle):
    bas = small_system("cc-pvdz", nw=2)
    fb = L.FockBuilder(bas)
    D1, d1 = density_for(oracle, bas, 10, seed=1)
    D2, d2 = density_for(oracle, bas, 10, seed=2)
    d_old = np.zeros_like(d1); f_old = np.zeros_like(d1)
    fb.fock_jk_inc(d1, d_old, f_old)
    f = fb.fock_jk_inc(d2, d_old, f_old)
    fo, _ = oracle.fock_jk(bas, d2)
    assert np.abs(f - fo).max() < 1e-9
    assert np.array_equal(d_old, d2)
    fb.close()


@pytest.mark.parametrize("cut", [1e-4, 1e-6])
def test_mixed_precision_energy(L, oracle, cut):
    """FP32 for quartets with bound below the cutoff, FP64 accumulation:
    |dE2| <= 1 uEh against the FP64 oracle, and the FP32 share is nonzero."""
    from openqp_amd.autograd import packed_trace
    bas = small_system("cc-pvdz", nw=4, seed=9)
    D, dp = density_for(oracle, bas, 20, seed=10)
    fo, _ = oracle.fock_jk(bas, dp)
    fb = L.FockBuilder(bas, precision=1, fp32_cutoff=cut)
    f = fb.fock_jk(dp)
    n32 = fb.last_stats.quartets_fp32
    dE = abs(packed_trace(dp, f, bas.nbf) - packed_trace(dp, fo, bas.nbf))
    _say("  mixed cut", cut, "fp32 quartets", n32, "of", fb.last_stats.quartets_computed, "dE2", dE,
         "max|df|", np.abs(f - fo).max())
    assert n32 > 0
    assert dE <= 1e-6
    fb.close()


@pytest.mark.parametrize("name,nw", [("cc-pvdz", 2), ("6-31g(d)", 3)])
def test_uhf_parity(L, oracle, name, nw):
    """nfocks = 2 through the C ABI (and routec_fock_jk) against the oracle's
    int2_urohf_data_t restatement: same screening, 1e-10 elementwise."""
    from openqp_amd import basis as B
    _say("UHF parity", name, nw)
    bas = small_system(name, nw=nw)
    S = oracle.int1e(bas)[0]
    ds = []
    for k, sd in ((5 * nw, 11), (5 * nw - 1, 12)):
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
    _say("  max |f - f_oracle| =", err, "quartets", fb.last_stats.quartets_computed, int(st[1]))
    assert err <= TOL
    assert fb.last_stats.quartets_computed == int(st[1])
    fb.close()


def _spin_densities(oracle, bas, na, nb, seeds=(11, 12)):
    from openqp_amd import basis as B
    S = oracle.int1e(bas)[0]
    ds = []
    for k, sd in ((na, seeds[0]), (nb, seeds[1])):
        rng = np.random.default_rng(sd)
        C = rng.standard_normal((bas.nbf, k))
        Lc = np.linalg.cholesky(C.T @ S @ C)
        C = C @ np.linalg.inv(Lc).T
        ds.append(B.pack_tri(C @ C.T))
    return ds


@pytest.mark.parametrize("basis,nw", [("cc-pvdz", 3), ("cc-pvtz", 1)])
def test_uhf_single_pass_kernel_modes(L, oracle, basis, nw):
    """The single-pass open-shell build (one digestion of D+ = Da + Db with
    Coulomb and exchange, one of D- = Da - Db with exchange only, from the
    same integrals) on every kernel family: wavefront-task kernel
    (OQPH_JK_ROWS=0), bra-row kernel with and without the LDS density rows
    (2, 3), grouped kernel on every class pair (OQPH_JKG=2).  Each against
    the oracle's int2_urohf_data_t restatement (1e-10) with the oracle's
    computed-quartet count."""
    _say("UHF single pass, kernel modes", basis, nw)
    bas = small_system(basis, nw=nw, seed=4)
    da, db = _spin_densities(oracle, bas, 5 * nw, 5 * nw - 1)
    Q = oracle.schwarz(bas)
    ra, rb, st = oracle.fock_jk_uhf(bas, da, db, se=0.8, sc=1.0, cutoff=5e-11, Q=Q)
    saved = {k: os.environ.get(k) for k in ("OQPH_JK_ROWS", "OQPH_JKG")}
    try:
        for rows, jkg in (("0", "0"), ("2", "1"), ("3", "1"), ("1", "2")):
            os.environ["OQPH_JK_ROWS"] = rows
            os.environ["OQPH_JKG"] = jkg
            fb = L.FockBuilder(bas)
            fa, fbb = fb.fock_jk_uhf(da, db, se=0.8)
            kinds = sorted({p["kernel"] for p in fb.last_profile()})
            err = max(np.abs(fa - ra).max(), np.abs(fbb - rb).max())
            _say("  rows", rows, "jkg", jkg, "max |f - f_oracle| =", err, "quartets",
                 fb.last_stats.quartets_computed, int(st[1]), "kernels", len(kinds))
            assert err <= TOL
            assert fb.last_stats.quartets_computed == int(st[1])
            fb.close()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_uhf_equal_spins_is_rhf(L, oracle):
    """Da = Db = D/2 makes D- = 0: both spin Fock matrices equal the RHF
    build of D (sc J[D] - se K[D]/2), elementwise to rounding.  (No
    screening: the open-shell screen uses max(|Da|, |Db|), half the RHF
    density bound here.)"""
    bas = small_system("cc-pvdz", nw=3, seed=6)
    D, dp = density_for(oracle, bas, 15, seed=8)
    fb = L.FockBuilder(bas, cutoff=1e-20)
    g = fb.fock_jk(dp, se=0.6, sc=0.9)
    fa, fbb = fb.fock_jk_uhf(0.5 * dp, 0.5 * dp, se=0.6, sc=0.9)
    _say("  equal spins: max |fa - g|", np.abs(fa - g).max(), "max |fb - g|", np.abs(fbb - g).max())
    assert np.abs(fa - g).max() <= 1e-12
    assert np.abs(fbb - g).max() <= 1e-12
    fb.close()


@pytest.mark.parametrize("basis,nw,nocc", [("cc-pvdz", 3, 15), ("6-31g(d)", 4, 20)])
def test_tiled_kernel_modes(L, oracle, basis, nw, nocc):
    """The tiled bra-row kernel (bras sharing their first shell: one ket read
    and one J_cd atomic per tile) against the oracle and against the
    segmented kernels: OQPH_JKT=0 (off, the default), 1 ((ps|ps)), 2 (every
    class pair with a tiled kernel).  Same G to 1e-12, same computed and
    screened counts; the open-shell pass and mixed precision as well."""
    from openqp_amd.autograd import packed_trace
    _say("tiled kernel modes", basis, nw)
    bas = small_system(basis, nw=nw, seed=17)
    D, dp = density_for(oracle, bas, nocc, seed=23)
    fo, so = oracle.fock_jk(bas, dp)
    da, db = _spin_densities(oracle, bas, nocc, nocc - 1, seeds=(31, 32))
    Q = oracle.schwarz(bas)
    ra, rb, su = oracle.fock_jk_uhf(bas, da, db, se=1.0, sc=1.0, cutoff=5e-11, Q=Q)
    out, kinds = {}, {}
    old = os.environ.get("OQPH_JKT")
    try:
        for mode in ("0", "1", "2"):
            os.environ["OQPH_JKT"] = mode
            fb = L.FockBuilder(bas)
            f = fb.fock_jk(dp)
            st = (fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
            kinds[mode] = {p["kernel"] for p in fb.last_profile()}
            fa, fbb = fb.fock_jk_uhf(da, db)
            nu = fb.last_stats.quartets_computed
            fb.set_precision(1, 1e-4)
            fm = fb.fock_jk(dp)
            out[mode] = (f, st, fa, fbb, nu, fm)
            fb.close()
    finally:
        if old is None:
            os.environ.pop("OQPH_JKT", None)
        else:
            os.environ["OQPH_JKT"] = old
    for mode in ("0", "1", "2"):
        f, st, fa, fbb, nu, fm = out[mode]
        err = np.abs(f - fo).max()
        eu = max(np.abs(fa - ra).max(), np.abs(fbb - rb).max())
        dE = abs(packed_trace(dp, fm, bas.nbf) - packed_trace(dp, fo, bas.nbf))
        _say("  OQPH_JKT", mode, "max |f - f_oracle| =", err, "UHF", eu, "mixed dE2", dE, "kernels", sorted(kinds[mode]))
        assert err <= TOL and eu <= TOL
        assert st[0] == int(so[1]) and nu == int(su[1])
        assert dE <= 1e-6
    for mode in ("1", "2"):
        assert np.abs(out["0"][0] - out[mode][0]).max() <= 1e-12
        assert out["0"][1] == out[mode][1]
    assert "k_jkt" not in kinds["0"]
    assert "k_jkt" in kinds["1"] and "k_jkt" in kinds["2"]


def test_large_d_classes_template_and_runtime(L, oracle):
    """(dp|dp), (dd|pp) and (dd|dp) run on the run-time-class kernel by
    default (more than 300 Cartesian components per quartet); OQPH_JK_HL_NG=0
    puts them back on their class templates.  Both against the oracle, same G
    to 1e-12 and the same quartet counts."""
    _say("large d classes: templates vs run-time class")
    bas = small_system("cc-pvdz", nw=3, seed=29)
    D, dp = density_for(oracle, bas, 15, seed=30)
    fo, so = oracle.fock_jk(bas, dp)
    out, kinds = {}, {}
    old = os.environ.get("OQPH_JK_HL_NG")
    try:
        for mode in ("0", None):
            if mode is None:
                os.environ.pop("OQPH_JK_HL_NG", None)
            else:
                os.environ["OQPH_JK_HL_NG"] = mode
            fb = L.FockBuilder(bas)
            out[mode] = (fb.fock_jk(dp), fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
            kinds[mode] = sorted({(p["kernel"], p["cls"]) for p in fb.last_profile()})
            fb.close()
    finally:
        if old is None:
            os.environ.pop("OQPH_JK_HL_NG", None)
        else:
            os.environ["OQPH_JK_HL_NG"] = old
    for mode in ("0", None):
        err = np.abs(out[mode][0] - fo).max()
        _say("  OQPH_JK_HL_NG", mode, "max |f - f_oracle| =", err)
        assert err <= TOL
        assert out[mode][1] == int(so[1])
    assert np.abs(out["0"][0] - out[None][0]).max() <= 1e-12
    assert out["0"][1:] == out[None][1:]
    assert ("k_jk", 2121) in kinds["0"] and ("k_jk", 2121) not in kinds[None]


def test_launch_order_modes(L, oracle):
    """OQPH_LAUNCH_ORDER=1 (default: launches in decreasing per-quartet cost)
    and 0 (class order) give the same G (1e-12), the same counts and the same
    2e gradient (1e-12); both against the oracle."""
    _say("launch order 1 vs 0")
    bas = small_system("cc-pvdz", nw=3, seed=31)
    D, dp = density_for(oracle, bas, 15, seed=32)
    fo, so = oracle.fock_jk(bas, dp)
    go, _ = oracle.grad2(bas, dp)
    out = {}
    old = os.environ.get("OQPH_LAUNCH_ORDER")
    try:
        for mode in ("1", "0"):
            os.environ["OQPH_LAUNCH_ORDER"] = mode
            fb = L.FockBuilder(bas)
            f = fb.fock_jk(dp)
            st = (fb.last_stats.quartets_computed, fb.last_stats.quartets_screened)
            g = fb.grad2(dp)
            out[mode] = (f, st, g)
            fb.close()
    finally:
        if old is None:
            os.environ.pop("OQPH_LAUNCH_ORDER", None)
        else:
            os.environ["OQPH_LAUNCH_ORDER"] = old
    assert np.abs(out["1"][0] - out["0"][0]).max() <= 1e-12
    assert out["1"][1] == out["0"][1] and out["1"][1][0] == so[1]
    assert np.abs(out["1"][0] - fo).max() <= 1e-10
    assert np.abs(out["1"][2] - out["0"][2]).max() <= 1e-12
    assert np.abs(out["1"][2] - go).max() <= 1e-8
