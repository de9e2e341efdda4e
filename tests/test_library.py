"""The C-ABI library loads and exports every entry point include/oqph.h
declares; the Route-C names match OpenQP's seam.  CPU only: no HIP calls."""
import ctypes
import os

import pytest

from helpers import header_functions, ROOT


def _lib_path():
    return os.path.join(ROOT, "openqp_amd", "_lib", "liboqph.so")


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "openqp_amd", "_lib", "liboqph.so")),
                    reason="liboqph not built")
def test_exports_match_header():
    lib = ctypes.CDLL(_lib_path())
    names = header_functions()
    assert "routec_fock_jk" in names and "routec_grad2" in names
    for n in names:
        assert hasattr(lib, n), n
    from openqp_amd import lib as L
    assert sorted(L.EXPORTS) == names


def test_routec_signature_matches_reference_seam():
    """routec_bridge.F90:33-61 binds (d,f,nbf,nfocks,se,sc,info) and
    (d,xyz,de,nbf,natm,hfscale,coulscale,info), all by reference."""
    txt = open(os.path.join(ROOT, "include", "oqph.h")).read()
    assert ("void routec_fock_jk(const double *d, double *f, const int *nbf, const int *nfocks, "
            "const double *se,") in txt
    assert ("void routec_grad2(const double *d, const double *xyz, double *de, const int *nbf, "
            "const int *natm,") in txt


def test_routec_grad2_mrsf_signature_matches_reference_seam():
    """routec_bridge.F90:61-75 binds routec_grad2_mrsf(d, p, spc, xyz, de, nbf,
    natm, hfscale, hfscale2, coulscale, spcscale, mrst, info), all by reference."""
    txt = " ".join(open(os.path.join(ROOT, "include", "oqph.h")).read().split())
    assert ("void routec_grad2_mrsf(const double *d, const double *p, const double *spc, const double *xyz, "
            "double *de, const int *nbf, const int *natm, const double *hfscale, const double *hfscale2, "
            "const double *coulscale, const double *spcscale, const int *mrst, int *info);") in txt


def test_routec_sig_signatures_match_reference_seam():
    """routec_sig.F90:31-57 binds routec_sig_init(nbf, mo_a, mo_b, fmo_a,
    fmo_b, nocca, noccb, kind) -> int, routec_sig_set_scale(s),
    routec_sig_iter(bvec_mo, nv_new, sigma_mo, info), routec_sig_free()."""
    txt = " ".join(open(os.path.join(ROOT, "include", "oqph.h")).read().split())
    assert ("int routec_sig_init(const int *nbf, const double *mo_a, const double *mo_b, const double *fmo_a, "
            "const double *fmo_b, const int *nocca, const int *noccb, const int *kind);") in txt
    assert "void routec_sig_set_scale(const double *s);" in txt
    assert "void routec_sig_iter(const double *bvec_mo, const int *nv_new, double *sigma_mo, int *info);" in txt
    assert "void routec_sig_free(void);" in txt


def test_product_does_not_import_oracle():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import openqp_amd, openqp_amd.lib, openqp_amd.molecules; "
            "assert not any(m == 'oracle' or m.startswith('oracle.') for m in sys.modules)" % ROOT)
    subprocess.check_call([sys.executable, "-c", code])


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "openqp_amd", "_lib", "liboqph.so")),
                    reason="liboqph not built")
def test_routec_sig_declines_without_session():
    """routec_sig_init declines (nonzero) with no default session and
    routec_sig_iter reports info != 0 before a successful init, so OpenQP keeps
    its native sigma path (routec_sig.F90:213-224); no device call is made."""
    lib = ctypes.CDLL(_lib_path())
    i = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    a = (ctypes.c_double * 4)()
    lib.routec_sig_init.argtypes = [ctypes.POINTER(i), dp, dp, dp, dp, ctypes.POINTER(i), ctypes.POINTER(i),
                                    ctypes.POINTER(i)]
    rc = lib.routec_sig_init(ctypes.byref(i(2)), a, a, a, a, ctypes.byref(i(2)), ctypes.byref(i(0)), ctypes.byref(i(1)))
    assert rc != 0
    lib.routec_sig_iter.argtypes = [dp, ctypes.POINTER(i), dp, ctypes.POINTER(i)]
    lib.routec_sig_iter.restype = None
    info = i(0)
    lib.routec_sig_iter(a, ctypes.byref(i(1)), a, ctypes.byref(info))
    assert info.value != 0


def test_routec_vxc_signature_matches_reference_seam():
    """routec_bridge.F90:43-50 binds routec_vxc(d, fxc, nbf, nfocks, eexc,
    totele, info), all by reference."""
    txt = " ".join(open(os.path.join(ROOT, "include", "oqph.h")).read().split())
    assert ("void routec_vxc(const double *d, double *fxc, const int *nbf, const int *nfocks, double *eexc, "
            "double *totele, int *info);") in txt


def test_routec_vxc_declines_without_session():
    """With no default session routec_vxc reports info = 1, so calc_jk_xc keeps
    its native grid path (scf_addons.F90:2104-2111); no device call is made."""
    lib = ctypes.CDLL(_lib_path())
    i = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    a = (ctypes.c_double * 4)()
    e, n = ctypes.c_double(0.0), ctypes.c_double(0.0)
    lib.routec_vxc.argtypes = [dp, dp, ctypes.POINTER(i), ctypes.POINTER(i), dp, dp, ctypes.POINTER(i)]
    lib.routec_vxc.restype = None
    info = i(0)
    lib.routec_vxc(a, a, ctypes.byref(i(2)), ctypes.byref(i(1)), ctypes.byref(e), ctypes.byref(n), ctypes.byref(info))
    assert info.value == 1


def test_xc_functional_table():
    """Exact-exchange fractions of the libxc compositions (libxc.F90:300-460)."""
    lib = ctypes.CDLL(_lib_path())
    lib.oqph_xc_functional_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int)]
    want = {b"b3lyp5": (0.2, 1), b"B3LYPV5": (0.2, 1), b"b3lypv1r": (0.2, 1), b"bhhlyp": (0.5, 1),
            b"blyp": (0.0, 1), b"svwn5": (0.0, 0), b"slater": (0.0, 0)}
    for name, (hf, gga) in want.items():
        h, g = ctypes.c_double(), ctypes.c_int()
        assert lib.oqph_xc_functional_info(name, ctypes.byref(h), ctypes.byref(g)) == 0
        assert (h.value, g.value) == (hf, gga), name
    assert lib.oqph_xc_functional_info(b"m06-2x", None, None) == 1
