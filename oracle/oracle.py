"""ctypes front end of the CPU FP64 oracle (oracle/oqo.c).

Test infrastructure only: imported by tests/, by the package smoke check and
by the CPU-baseline leg of bench.py -- always as the checker or the reported
CPU baseline, never on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboqo.so")
_lib = None

_pi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_BAS = [C.c_int, C.c_int, C.c_int, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.oqo_boys.argtypes = [C.c_int, C.c_double, _pd]
        L.oqo_c2s.argtypes = [C.c_int, _pd]
        L.oqo_cart_order.argtypes = [C.c_int, _pi]
        L.oqo_schwarz.argtypes = _BAS + [_pd]
        L.oqo_schwarz_mu.argtypes = _BAS + [C.c_double, _pd]
        L.oqo_eri_quartet.argtypes = _BAS + [C.c_int] * 4 + [C.c_double, _pd]
        L.oqo_eri_quartet.restype = C.c_int
        L.oqo_eri_quartet_mu.argtypes = _BAS + [C.c_int] * 4 + [C.c_double, C.c_double, _pd]
        L.oqo_eri_quartet_mu.restype = C.c_int
        _pl = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.oqo_pair_map.argtypes = [C.c_int, _pi, _pd, _pl]
        L.oqo_fock_jk_ex.argtypes = _BAS + [_pd, C.c_void_p, _pd, _pd, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double, _pl, C.c_int, C.c_int]
        L.oqo_fock_jk.argtypes = _BAS + [_pd, _pd, _pd, C.c_double, C.c_double, C.c_double,
                                         np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.oqo_fock_jk_part.argtypes = L.oqo_fock_jk.argtypes + [C.c_int, C.c_int]
        L.oqo_fock_jk_uhf.argtypes = L.oqo_fock_jk.argtypes
        L.oqo_mrsf_int2.argtypes = _BAS + [_pd, C.c_int, C.c_int, C.c_int, _pd, _pd, C.c_double, C.c_double,
                                           C.c_double, _pl]
        L.oqo_eri_tensor.argtypes = _BAS + [_pd]
        L.oqo_int1e.argtypes = _BAS + [_pd, _pd, _pd, _pd]
        L.oqo_grad2_mrsf.argtypes = _BAS + [_pd, _pd, _pd, _pd, _pd, C.c_double, C.c_double, C.c_double, _pd,
                                            C.c_int, C.c_double, _pl]
        if hasattr(L, "oqo_grad2"):
            L.oqo_grad2.argtypes = _BAS + [_pd, _pd, _pd, C.c_double, C.c_double, C.c_double,
                                           np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        _lib = L
    return _lib


def boys(mmax: int, T: float) -> np.ndarray:
    F = np.zeros(mmax + 1)
    lib().oqo_boys(mmax, float(T), F)
    return F


def c2s(l: int) -> np.ndarray:
    nc, ns = (l + 1) * (l + 2) // 2, 2 * l + 1
    out = np.zeros(nc * ns)
    lib().oqo_c2s(l, out)
    return out.reshape(nc, ns)


def cart_order(l: int) -> np.ndarray:
    nc = (l + 1) * (l + 2) // 2
    out = np.zeros(nc * 3, dtype=np.int32)
    lib().oqo_cart_order(l, out)
    return out.reshape(nc, 3)


def schwarz(bas, mu=None) -> np.ndarray:
    """Schwarz table sqrt(max|(ab|ab)|); mu: of the erf(mu r)/r integrals."""
    Q = np.zeros(bas.nsh * bas.nsh)
    if mu is None:
        lib().oqo_schwarz(*bas.carray(), Q)
    else:
        lib().oqo_schwarz_mu(*bas.carray(), 1.0 / (mu * mu), Q)
    return Q.reshape(bas.nsh, bas.nsh)


def eri_quartet(bas, i, j, k, l, cutoff=0.0, mu=None) -> np.ndarray:
    """Final-representation ERI block; mu: erf(mu r)/r operator instead of 1/r."""
    n = bas.naos()
    out = np.zeros(int(n[i] * n[j] * n[k] * n[l]))
    if mu is None:
        lib().oqo_eri_quartet(*bas.carray(), i, j, k, l, cutoff, out)
    else:
        lib().oqo_eri_quartet_mu(*bas.carray(), i, j, k, l, cutoff, 1.0 / (mu * mu), out)
    return out.reshape(n[i], n[j], n[k], n[l])


def pair_map(bas, Q) -> np.ndarray:
    """1-based position ij_pair of every shell pair i >= j (index i(i+1)/2 + j) in
    OpenQP's cost-sorted shell-pair map (int2_build_shell_pair_map)."""
    out = np.zeros(bas.nsh * (bas.nsh + 1) // 2, dtype=np.int64)
    lib().oqo_pair_map(bas.nsh, np.ascontiguousarray(bas.am, np.int32), np.ascontiguousarray(Q.reshape(-1)), out)
    return out


def fock_jk_general(bas, d, nfocks=1, se=1.0, sc=1.0, cutoff=5e-11, Q=None, Qmap=None, mu=None, rank=0, size=1):
    """Packed f (nfocks * ntri) and stats for 1/r (mu None) or erf(mu r)/r, on
    the share (rank, size); Q screens, Qmap orders the shell-pair map."""
    if Q is None:
        Q = schwarz(bas, mu)
    f = np.zeros(nfocks * bas.ntri)
    stats = np.zeros(3, dtype=np.int64)
    qm = None if Qmap is None else np.ascontiguousarray(Qmap.reshape(-1), np.float64)
    lib().oqo_fock_jk_ex(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)),
                         None if qm is None else qm.ctypes.data, np.ascontiguousarray(d, np.float64), f, nfocks,
                         se, sc, cutoff, 0.0 if mu is None else 1.0 / (mu * mu), stats, rank, size)
    return f, stats


def fock_jk_cam(bas, d, alpha, beta, mu, alpha_coulomb=1.0, beta_coulomb=0.0, nfocks=1, cutoff=5e-11):
    """CAM build of int2_run_cam (int2.F90:538-584): 1/r pass (Coulomb
    alpha_coulomb, exchange alpha) plus erf(mu r)/r pass (Coulomb beta_coulomb,
    exchange beta) screened with the attenuated bounds."""
    f1, s1 = fock_jk_general(bas, d, nfocks, alpha, alpha_coulomb, cutoff)
    f2, s2 = fock_jk_general(bas, d, nfocks, beta, beta_coulomb, cutoff, mu=mu)
    return f1 + f2, s1 + s2


def eri_tensor(bas) -> np.ndarray:
    nbf = bas.nbf
    out = np.zeros(nbf ** 4)
    lib().oqo_eri_tensor(*bas.carray(), out)
    return out.reshape(nbf, nbf, nbf, nbf)


def int1e(bas):
    nbf = bas.nbf
    S, T, V = (np.zeros(nbf * nbf) for _ in range(3))
    lib().oqo_int1e(*bas.carray(), np.ascontiguousarray(bas.zn, np.float64), S, T, V)
    return S.reshape(nbf, nbf), T.reshape(nbf, nbf), V.reshape(nbf, nbf)


def fock_jk(bas, d_packed, se=1.0, sc=1.0, cutoff=5e-11, Q=None, rank=0, size=1):
    """Packed RHF 2e Fock f (scaled like fock_jk) and stats (nschwz, quartets, ints).
    rank/size restrict the build to OpenQP's round-robin share of shell pairs."""
    if Q is None:
        Q = schwarz(bas)
    f = np.zeros(bas.ntri)
    stats = np.zeros(3, dtype=np.int64)
    lib().oqo_fock_jk_part(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)),
                           np.ascontiguousarray(d_packed, np.float64), f, se, sc, cutoff, stats, rank, size)
    return f, stats


def fock_jk_uhf(bas, d_alpha, d_beta, se=1.0, sc=1.0, cutoff=5e-11, Q=None):
    """Packed UHF/ROHF 2e Fock matrices (f_alpha, f_beta) for spin densities
    (int2_urohf_data_t_update semantics) and stats."""
    if Q is None:
        Q = schwarz(bas)
    d = np.concatenate([np.asarray(d_alpha, np.float64), np.asarray(d_beta, np.float64)])
    f = np.zeros(2 * bas.ntri)
    stats = np.zeros(3, dtype=np.int64)
    lib().oqo_fock_jk_uhf(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)), d, f, se, sc, cutoff, stats)
    return f[:bas.ntri], f[bas.ntri:], stats


def mrsf_int2(bas, d3, ncoul=4, se=1.0, sc=1.0, cutoff=5e-11, Q=None):
    """MRSF response integral pass (int2_mrsf_data_t): d3 of shape
    (nbf, nbf, nmat, nf) is the reference's Fortran d3(nf, nmat, nbf, nbf)
    read in C order, i.e. d3[nu, mu, c, v] = d3_F(v, c, mu, nu); returns f3 of
    the same shape and the stats."""
    if Q is None:
        Q = schwarz(bas)
    d3 = np.ascontiguousarray(d3, np.float64)
    nbf2, nbf1, nmat, nf = d3.shape
    f3 = np.zeros_like(d3)
    stats = np.zeros(3, dtype=np.int64)
    lib().oqo_mrsf_int2(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)), nf, nmat, ncoul, d3, f3, se, sc,
                        cutoff, stats)
    return f3, stats


def grad2_mrsf(bas, d, p, spc, hfscale=1.0, hfscale2=1.0, coulscale=1.0, spcscale=(1.0, 1.0, 1.0), mrst=1,
               cutoff=1e-10, Q=None):
    """MRSF 2e gradient (natm, 3): d, p (2, nbf, nbf) raw alpha/beta densities
    as numpy [spin, j, i] = Fortran (i, j, spin); spc (nbf, nbf, 7) as numpy
    [j, i, c] = Fortran spc(c, i, j)."""
    if Q is None:
        Q = schwarz(bas)
    de = np.zeros(bas.natm * 3)
    stats = np.zeros(3, dtype=np.int64)
    lib().oqo_grad2_mrsf(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)), np.ascontiguousarray(d, np.float64),
                         np.ascontiguousarray(p, np.float64), np.ascontiguousarray(spc, np.float64), de, hfscale,
                         hfscale2, coulscale, np.ascontiguousarray(spcscale, np.float64), mrst, cutoff, stats)
    return de.reshape(bas.natm, 3), stats


def grad2(bas, d_packed, hfscale=1.0, coulscale=1.0, cutoff=1e-10, Q=None):
    """dE2/dR (natm, 3) at fixed density, Coulomb + hfscale * exchange."""
    if Q is None:
        Q = schwarz(bas)
    de = np.zeros(bas.natm * 3)
    stats = np.zeros(3, dtype=np.int64)
    lib().oqo_grad2(*bas.carray(), np.ascontiguousarray(Q.reshape(-1)),
                    np.ascontiguousarray(d_packed, np.float64), de, hfscale, coulscale, cutoff, stats)
    return de.reshape(bas.natm, 3), stats
