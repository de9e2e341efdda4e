"""ctypes binding of liboqph (include/oqph.h) and the host-side mirror of
OpenQP's Fock-build interface.

``FockBuilder`` plays the role of ``int2_compute_t`` + ``fock_jk``
(source/integrals/int2.F90:137-185, source/scf_addons.F90:1063-1214): it is
initialised with a basis, computes the Schwarz table, and returns the packed
2e Fock contribution for a packed density.  The device library is required:
every call goes to liboqph and a missing library is an error, never a CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .basis import Basis

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liboqph.so")
_lib = None

_pi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

EXPORTS = ["routec_fock_jk", "routec_grad2", "oqph_session_create", "oqph_session_destroy",
           "oqph_session_set_basis", "oqph_session_set_geometry", "oqph_session_set_thresholds",
           "oqph_session_set_precision", "oqph_set_default_session", "oqph_fock_jk", "oqph_fock_jk_inc",
           "oqph_grad2", "oqph_fock_jk_device", "oqph_schwarz", "oqph_session_nbf", "oqph_last_error",
           "oqph_session_set_profiling", "oqph_last_profile", "oqph_int1e", "oqph_grad1e",
           "oqph_session_prepare", "oqph_session_set_share", "oqph_fock_jk_cam", "oqph_kernel_resources",
           "oqph_fock_jk_noalloc", "oqph_last_counters", "oqph_session_set_grad_precision", "oqph_mrsf_int2",
           "routec_sig_init", "routec_sig_set_scale", "routec_sig_iter", "routec_sig_free", "oqph_grad2_mrsf",
           "routec_grad2_mrsf", "routec_vxc", "oqph_session_set_xc", "oqph_xc_functional_info", "oqph_vxc",
           "oqph_xc_grid", "oqph_xc_grid_stats"]


class OqphError(RuntimeError):
    pass


class Stats(C.Structure):
    _fields_ = [("quartets_computed", C.c_longlong), ("quartets_screened", C.c_longlong),
                ("quartets_fp32", C.c_longlong), ("kernel_ms", C.c_double),
                ("model_flops", C.c_double), ("prim_quartets", C.c_longlong)]


def load(path: str | None = None):
    """Load liboqph.so (the in-tree build, or the file $OQPH_LIB names); raises
    if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("OQPH_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise OqphError(f"liboqph not built: {p} missing (run openqp_amd/build.py)")
    L = C.CDLL(p)
    vp = C.c_void_p
    L.oqph_session_create.argtypes = [C.POINTER(vp), C.c_int]
    L.oqph_session_destroy.argtypes = [vp]
    L.oqph_session_set_basis.argtypes = [vp, C.c_int, _pi, _pi, _pi, _pi, _pd, _pd]
    L.oqph_session_set_geometry.argtypes = [vp, C.c_int, _pd, _pd]
    L.oqph_session_set_thresholds.argtypes = [vp, C.c_double, C.c_double]
    L.oqph_session_set_precision.argtypes = [vp, C.c_int, C.c_double]
    L.oqph_session_set_grad_precision.argtypes = [vp, C.c_int, C.c_double]
    L.oqph_set_default_session.argtypes = [vp]
    L.oqph_fock_jk.argtypes = [vp, _pd, _pd, C.c_int, C.c_double, C.c_double, C.POINTER(Stats)]
    L.oqph_fock_jk_inc.argtypes = [vp, _pd, _pd, _pd, _pd, C.c_int, C.c_double, C.c_double, C.POINTER(Stats)]
    L.oqph_grad2.argtypes = [vp, _pd, _pd, C.c_double, C.c_double, C.POINTER(Stats)]
    L.oqph_fock_jk_device.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                      C.POINTER(Stats), C.c_void_p]
    L.oqph_fock_jk_noalloc.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.POINTER(Stats),
                                       C.c_void_p]
    L.oqph_schwarz.argtypes = [vp, _pd]
    L.oqph_mrsf_int2.argtypes = [vp, C.c_int, C.c_int, C.c_int, _pd, _pd, C.c_double, C.c_double,
                                 C.POINTER(Stats)]
    L.oqph_session_prepare.argtypes = [vp]
    L.oqph_session_set_share.argtypes = [vp, C.c_int, C.c_int]
    L.oqph_fock_jk_cam.argtypes = [vp, _pd, _pd, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.POINTER(Stats)]
    L.oqph_kernel_resources.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _pi]
    L.oqph_session_nbf.argtypes = [vp]
    L.oqph_last_error.restype = C.c_char_p
    L.oqph_session_set_profiling.argtypes = [vp, C.c_int]
    _pl = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    L.oqph_last_profile.argtypes = [vp, C.c_int, _pi, _pd, _pd, _pl, _pl]
    L.oqph_last_counters.argtypes = [vp, C.c_int, np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")]
    L.oqph_int1e.argtypes = [C.c_int, _pi, _pi, _pi, _pi, _pd, _pd, C.c_int, _pd, _pd, C.c_void_p, C.c_void_p,
                             C.c_void_p]
    L.oqph_grad1e.argtypes = [C.c_int, _pi, _pi, _pi, _pi, _pd, _pd, C.c_int, _pd, _pd, C.c_void_p, C.c_void_p,
                              C.c_void_p]
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    L.routec_fock_jk.argtypes = [dp, dp, ip, ip, dp, dp, ip]
    L.routec_grad2.argtypes = [dp, dp, dp, ip, ip, dp, dp, ip]
    L.routec_sig_init.argtypes = [ip, dp, dp, dp, dp, ip, ip, ip]
    L.oqph_grad2_mrsf.argtypes = [vp, _pd, _pd, _pd, _pd, C.c_double, C.c_double, C.c_double, _pd, C.c_int,
                                  C.POINTER(Stats)]
    L.routec_grad2_mrsf.argtypes = [dp, dp, dp, dp, dp, ip, ip, dp, dp, dp, dp, ip, ip]
    L.routec_grad2_mrsf.restype = None
    L.routec_sig_set_scale.argtypes = [dp]
    L.routec_sig_iter.argtypes = [dp, ip, dp, ip]
    L.routec_sig_set_scale.restype = None
    L.routec_sig_iter.restype = None
    L.routec_sig_free.restype = None
    L.routec_vxc.argtypes = [dp, dp, ip, ip, dp, dp, ip]
    L.routec_vxc.restype = None
    L.oqph_session_set_xc.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
    L.oqph_xc_functional_info.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.oqph_vxc.argtypes = [vp, _pd, C.c_int, _pd, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.oqph_xc_grid.argtypes = [vp, C.c_void_p, C.c_longlong]
    L.oqph_xc_grid.restype = C.c_longlong
    L.oqph_xc_grid_stats.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        raise OqphError(f"{what} failed (code {rc}): {load().oqph_last_error().decode()}")


class FockBuilder:
    """Device J/K builder for one basis + geometry (one liboqph session)."""

    def __init__(self, bas: Basis, cutoff: float = 5e-11, grad_cutoff: float = 1e-10, device: int = 0,
                 precision: int = 0, fp32_cutoff: float = 0.0):
        L = load()
        self.bas = bas
        self._s = C.c_void_p()
        _check(L.oqph_session_create(C.byref(self._s), device), "session_create")
        _check(L.oqph_session_set_basis(self._s, bas.nsh, np.ascontiguousarray(bas.am, np.int32),
                                        np.ascontiguousarray(bas.harmonic, np.int32),
                                        np.ascontiguousarray(bas.atom, np.int32),
                                        np.ascontiguousarray(bas.nprim, np.int32),
                                        np.ascontiguousarray(bas.ex, np.float64),
                                        np.ascontiguousarray(bas.cc, np.float64)), "set_basis")
        self.set_geometry(bas.xyz)
        _check(L.oqph_session_set_thresholds(self._s, cutoff, grad_cutoff), "set_thresholds")
        _check(L.oqph_session_set_precision(self._s, precision, fp32_cutoff), "set_precision")
        self.last_stats = Stats()

    def set_geometry(self, xyz):
        self.bas = self.bas.with_coords(np.asarray(xyz, np.float64).reshape(-1, 3))
        _check(load().oqph_session_set_geometry(self._s, self.bas.natm, np.ascontiguousarray(self.bas.xyz.reshape(-1)),
                                                np.ascontiguousarray(self.bas.zn, np.float64)), "set_geometry")

    def set_precision(self, policy: int, fp32_cutoff: float = 0.0):
        _check(load().oqph_session_set_precision(self._s, policy, fp32_cutoff), "set_precision")

    def set_grad_precision(self, policy: int, fp32_cutoff: float = 1e-4):
        """2e-gradient precision: 0 = FP64; 1 = quartets with Q_ij Q_kl < fp32_cutoff in FP32."""
        _check(load().oqph_session_set_grad_precision(self._s, policy, fp32_cutoff), "set_grad_precision")

    def grad2_mrsf(self, d, p, spc, hfscale: float = 1.0, hfscale2: float = 1.0, coulscale: float = 1.0,
                   spcscale=(1.0, 1.0, 1.0), mrst: int = 1) -> np.ndarray:
        """MRSF 2e gradient (natm, 3): d, p as numpy (2, nbf, nbf) = Fortran
        (nbf, nbf, 2) raw alpha/beta densities, spc as numpy (nbf, nbf, 7) =
        Fortran spc(7, nbf, nbf)."""
        de = np.zeros(3 * self.bas.natm)
        _check(load().oqph_grad2_mrsf(self._s, np.ascontiguousarray(d, np.float64), np.ascontiguousarray(p, np.float64),
                                      np.ascontiguousarray(spc, np.float64), de, hfscale, hfscale2, coulscale,
                                      np.ascontiguousarray(spcscale, np.float64), mrst, C.byref(self.last_stats)),
               "grad2_mrsf")
        return de.reshape(self.bas.natm, 3)

    def mrsf_int2(self, d3, ncoul: int = 4, se: float = 1.0, sc: float = 1.0) -> np.ndarray:
        """MRSF response integral pass (int2_mrsf_data_t): d3 of shape (nbf, nbf,
        nmat, nf) is the reference's Fortran d3(nf, nmat, nbf, nbf) in C order;
        returns f3 of the same shape."""
        d3 = np.ascontiguousarray(d3, np.float64)
        nbf2, nbf1, nmat, nf = d3.shape
        if nbf1 != self.bas.nbf or nbf2 != self.bas.nbf:
            raise OqphError(f"mrsf_int2: d3 has {nbf1}x{nbf2} AOs, the basis {self.bas.nbf}")
        f3 = np.empty_like(d3)
        _check(load().oqph_mrsf_int2(self._s, nf, nmat, ncoul, d3, f3, se, sc, C.byref(self.last_stats)),
               "mrsf_int2")
        return f3

    def set_share(self, rank: int, size: int):
        """Restrict builds to OpenQP MPI rank `rank` of `size` (int2.F90:759-761)."""
        _check(load().oqph_session_set_share(self._s, rank, size), "set_share")

    def prepare(self):
        """Build the geometry-dependent device state (and the rank share) now."""
        _check(load().oqph_session_prepare(self._s), "prepare")

    def fock_jk_cam(self, d_packed, alpha: float, beta: float, mu: float, alpha_coulomb: float = 1.0,
                    beta_coulomb: float = 0.0, nfocks: int = 1) -> np.ndarray:
        """Packed CAM G = G[1/r; alpha_coulomb, alpha] + G[erf(mu r)/r; beta_coulomb, beta] (int2_run_cam)."""
        d = np.ascontiguousarray(d_packed, np.float64)
        f = np.zeros_like(d)
        _check(load().oqph_fock_jk_cam(self._s, d, f, nfocks, alpha, beta, mu, alpha_coulomb, beta_coulomb,
                                       C.byref(self.last_stats)), "fock_jk_cam")
        return f

    def make_default(self):
        _check(load().oqph_set_default_session(self._s), "set_default_session")

    def set_profiling(self, on: bool = True):
        _check(load().oqph_session_set_profiling(self._s, 1 if on else 0), "set_profiling")

    def last_profile(self):
        """Per-launch (class code, ms, model FLOPs, quartets, primitive quartets) of the last build.

        cls = la*1000 + lb*100 + lc*10 + ld; kernel is "k_jkb" (bra-row kernel), "k_jkg"
        (grouped bra-row kernel), "k_jkt" (its tiled form: bras sharing their first
        shell), "k_jk" (wavefront tasks) or "k_jkh" (run-time class,
        the classes with an f shell); a launch of several run-time classes has
        kernel "k_jkh_multi" and cls 40000 + its number of classes."""
        n = 64
        cls = np.zeros(n, np.int32); ms = np.zeros(n); fl = np.zeros(n)
        q = np.zeros(n, np.int64); pq = np.zeros(n, np.int64)
        m = load().oqph_last_profile(self._s, n, cls, ms, fl, q, pq)
        kern = lambda c: ("k_jkt" if c >= 60000 else "k_jkh_multi" if c >= 40000 else "k_jkg" if c >= 30000
                          else "k_jkh" if c >= 20000 else "k_jkb" if c >= 10000 else "k_jk")
        ctr = np.zeros(8 * n, np.uint64)
        load().oqph_last_counters(self._s, n, ctr)
        return [dict(cls=int(cls[i]) if 40000 <= cls[i] < 50000 else int(cls[i]) % 10000, kernel=kern(int(cls[i])),
                     ms=float(ms[i]),
                     flops=float(fl[i]), quartets=int(q[i]), prims=int(pq[i]), prims_eval=int(ctr[8 * i + 5]),
                     screened=int(ctr[8 * i + 1]), slots=int(ctr[8 * i + 6]))
                for i in range(max(m, 0))]

    def schwarz(self) -> np.ndarray:
        q = np.zeros(self.bas.nsh * self.bas.nsh)
        _check(load().oqph_schwarz(self._s, q), "schwarz")
        return q.reshape(self.bas.nsh, self.bas.nsh)

    def fock_jk(self, d_packed, se: float = 1.0, sc: float = 1.0) -> np.ndarray:
        """Packed G = sc J - se K/2 for a packed total density (fock_jk semantics)."""
        d = np.ascontiguousarray(d_packed, np.float64)
        f = np.zeros_like(d)
        _check(load().oqph_fock_jk(self._s, d, f, 1, se, sc, C.byref(self.last_stats)), "fock_jk")
        return f

    def set_xc(self, functional: str, nrad: int = 96, ntheta: int = 15):
        """Select the XC functional and the grid (nrad radial shells, ntheta x
        2 ntheta angular points per shell) of this session's Vxc builds."""
        _check(load().oqph_session_set_xc(self._s, functional.encode(), nrad, ntheta), "set_xc")

    def vxc(self, d, nfocks: int = 1):
        """Packed Vxc per spin, E_xc and the grid electron count for the packed
        density d (nfocks = 1: total density; 2: alpha then beta)."""
        d = np.ascontiguousarray(d, np.float64).reshape(-1)
        f = np.zeros_like(d)
        exc, nel = C.c_double(), C.c_double()
        _check(load().oqph_vxc(self._s, d, nfocks, f, C.byref(exc), C.byref(nel)), "vxc")
        n = d.size // nfocks
        return [f[i * n:(i + 1) * n] for i in range(nfocks)], exc.value, nel.value

    def xc_grid(self) -> np.ndarray:
        """The session's XC grid: (npoints, 4) rows x, y, z (bohr), weight."""
        L = load()
        n = L.oqph_xc_grid(self._s, None, 0)
        if n < 0:
            raise OqphError(f"xc_grid failed (code {n}): {L.oqph_last_error().decode()}")
        out = np.zeros((n, 4))
        L.oqph_xc_grid(self._s, out.ctypes.data_as(C.c_void_p), n)
        return out

    def xc_grid_stats(self) -> dict:
        """Points, batches and GEMM work (sum of points x basis functions^k) of a Vxc build."""
        out = np.zeros(5, dtype=np.int64)
        _check(load().oqph_xc_grid_stats(self._s, out), "xc_grid_stats")
        return {"points": int(out[0]), "batches": int(out[1]), "sum_np_nb": int(out[2]), "sum_np_nb2": int(out[3]),
                "max_nb": int(out[4])}

    def fock_jk_uhf(self, d_alpha, d_beta, se: float = 1.0, sc: float = 1.0):
        """Packed UHF/ROHF (F_alpha, F_beta) = sc J[Da + Db] - se K[Ds] (fock_jk with nfocks = 2)."""
        d = np.ascontiguousarray(np.concatenate([np.asarray(d_alpha, np.float64), np.asarray(d_beta, np.float64)]))
        f = np.zeros_like(d)
        _check(load().oqph_fock_jk(self._s, d, f, 2, se, sc, C.byref(self.last_stats)), "fock_jk (nfocks=2)")
        n = d.size // 2
        return f[:n], f[n:]

    def fock_jk_inc(self, d, d_old, f_old, se: float = 1.0, sc: float = 1.0):
        """Incremental build: returns f = G[d - d_old] + f_old and updates d_old, f_old in place."""
        d = np.ascontiguousarray(d, np.float64)
        f = np.zeros_like(d)
        _check(load().oqph_fock_jk_inc(self._s, d, d_old, f, f_old, 1, se, sc, C.byref(self.last_stats)),
               "fock_jk_inc")
        return f

    def fock_jk_device(self, d_dev_ptr: int, f_dev_ptr: int, se: float = 1.0, sc: float = 1.0,
                       stream: int = 0) -> Stats:
        st = Stats()
        _check(load().oqph_fock_jk_device(self._s, C.c_void_p(d_dev_ptr), C.c_void_p(f_dev_ptr), 1, se, sc,
                                          C.byref(st), C.c_void_p(stream)), "fock_jk_device")
        return st

    def fock_jk_noalloc(self, d_dev_ptr: int, f_dev_ptr: int, se: float = 1.0, sc: float = 1.0,
                        stream: int = 0) -> Stats:
        """RHF build through the no-allocation entry point (the session must be
        prepared: ``prepare()``)."""
        st = Stats()
        _check(load().oqph_fock_jk_noalloc(self._s, C.c_void_p(d_dev_ptr), C.c_void_p(f_dev_ptr), se, sc,
                                           C.byref(st), C.c_void_p(stream)), "fock_jk_noalloc")
        return st

    def grad2(self, d_packed, hfscale: float = 1.0, coulscale: float = 1.0) -> np.ndarray:
        d = np.ascontiguousarray(d_packed, np.float64)
        de = np.zeros(self.bas.natm * 3)
        _check(load().oqph_grad2(self._s, d, de, hfscale, coulscale, C.byref(self.last_stats)), "grad2")
        return de.reshape(self.bas.natm, 3)

    def close(self):
        if self._s:
            load().oqph_session_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_resources(kind: int, la: int, lb: int, lc: int, ld: int, pure: bool):
    """(private bytes per lane, static LDS, registers, largest block) of a class-templated kernel, or None."""
    out = np.zeros(4, np.int32)
    rc = load().oqph_kernel_resources(kind, la, lb, lc, ld, 1 if pure else 0, out)
    return None if rc != 0 else tuple(int(x) for x in out)


def int1e(bas: Basis):
    """S, T, V (nbf x nbf) in OpenQP's AO representation (host code in liboqph)."""
    L = load()
    n = bas.nbf
    S, T, V = (np.zeros((n, n)) for _ in range(3))
    rc = L.oqph_int1e(bas.nsh, np.ascontiguousarray(bas.am, np.int32), np.ascontiguousarray(bas.harmonic, np.int32),
                      np.ascontiguousarray(bas.atom, np.int32), np.ascontiguousarray(bas.nprim, np.int32),
                      np.ascontiguousarray(bas.ex, np.float64), np.ascontiguousarray(bas.cc, np.float64), bas.natm,
                      np.ascontiguousarray(bas.xyz.reshape(-1)), np.ascontiguousarray(bas.zn, np.float64),
                      S.ctypes.data, T.ctypes.data, V.ctypes.data)
    _check(rc, "int1e")
    return S, T, V


def grad1e(bas: Basis, D: np.ndarray, W: np.ndarray) -> np.ndarray:
    """One-electron gradient terms (natm x 3): sum D dH/dR - sum W dS/dR (host code in liboqph)."""
    L = load()
    g = np.zeros(3 * bas.natm)
    Dm = np.ascontiguousarray(D, np.float64)
    Wm = np.ascontiguousarray(W, np.float64)
    rc = L.oqph_grad1e(bas.nsh, np.ascontiguousarray(bas.am, np.int32), np.ascontiguousarray(bas.harmonic, np.int32),
                       np.ascontiguousarray(bas.atom, np.int32), np.ascontiguousarray(bas.nprim, np.int32),
                       np.ascontiguousarray(bas.ex, np.float64), np.ascontiguousarray(bas.cc, np.float64), bas.natm,
                       np.ascontiguousarray(bas.xyz.reshape(-1)), np.ascontiguousarray(bas.zn, np.float64),
                       Dm.ctypes.data, Wm.ctypes.data, g.ctypes.data)
    _check(rc, "grad1e")
    return g.reshape(bas.natm, 3)


def routec_grad2(d_packed, xyz_bohr, nbf: int, hfscale: float = 1.0, coulscale: float = 1.0):
    """Call the Route-C gradient symbol exactly as OpenQP's seam does (by reference,
    ``xyz(3, natm)`` in bohr); returns (de (natm, 3), info)."""
    L = load()
    d = np.ascontiguousarray(d_packed, np.float64)
    xyz = np.ascontiguousarray(np.asarray(xyz_bohr, np.float64).reshape(-1, 3))
    natm = xyz.shape[0]
    de = np.zeros(3 * natm)
    nb, na, info = C.c_int(nbf), C.c_int(natm), C.c_int(-1)
    hs, cs = C.c_double(hfscale), C.c_double(coulscale)
    dp = C.POINTER(C.c_double)
    L.routec_grad2(d.ctypes.data_as(dp), xyz.ctypes.data_as(dp), de.ctypes.data_as(dp), C.byref(nb), C.byref(na),
                   C.byref(hs), C.byref(cs), C.byref(info))
    return de.reshape(natm, 3), info.value


def routec_fock_jk(d_packed, nbf: int, se: float = 1.0, sc: float = 1.0):
    """Call the Route-C symbol exactly as OpenQP's seam does (by reference); returns (f, info)."""
    L = load()
    d = np.ascontiguousarray(d_packed, np.float64)
    f = np.zeros_like(d)
    nb, nf, info = C.c_int(nbf), C.c_int(1), C.c_int(-1)
    se_, sc_ = C.c_double(se), C.c_double(sc)
    L.routec_fock_jk(d.ctypes.data_as(C.POINTER(C.c_double)), f.ctypes.data_as(C.POINTER(C.c_double)),
                     C.byref(nb), C.byref(nf), C.byref(se_), C.byref(sc_), C.byref(info))
    return f, info.value


class SigmaSession:
    """The MRSF sigma seam (routec_sig_init / set_scale / iter / free) called by
    reference exactly as OpenQP's routec_sig module does; uses the default
    session (oqph_set_default_session) for the basis and geometry."""

    def __init__(self, mo_a, mo_b, fmo_a, fmo_b, nocca: int, noccb: int, kind: int, scale: float = 1.0):
        L = load()
        self.nbf = int(np.asarray(mo_a).shape[0])
        self.nocca, self.noccb = nocca, noccb
        dp = C.POINTER(C.c_double)
        f = [np.asfortranarray(x, np.float64) for x in (mo_a, mo_b, fmo_a, fmo_b)]
        rc = L.routec_sig_init(C.byref(C.c_int(self.nbf)), *[x.ctypes.data_as(dp) for x in f],
                               C.byref(C.c_int(nocca)), C.byref(C.c_int(noccb)), C.byref(C.c_int(kind)))
        if rc != 0:
            raise OqphError(f"routec_sig_init failed (code {rc}): {L.oqph_last_error().decode()}")
        L.routec_sig_set_scale(C.byref(C.c_double(scale)))

    def iter(self, bvec) -> np.ndarray:
        """(A - B) X for the columns of bvec (nocca (nbf - noccb) x nv)."""
        L = load()
        b = np.asfortranarray(np.asarray(bvec, np.float64).reshape(self.nocca * (self.nbf - self.noccb), -1))
        out = np.zeros_like(b, order="F")
        info = C.c_int(-1)
        dp = C.POINTER(C.c_double)
        L.routec_sig_iter(b.ctypes.data_as(dp), C.byref(C.c_int(b.shape[1])), out.ctypes.data_as(dp), C.byref(info))
        if info.value != 0:
            raise OqphError(f"routec_sig_iter failed (info {info.value}): {L.oqph_last_error().decode()}")
        return out

    def close(self):
        load().routec_sig_free()


def xc_functional_info(functional: str):
    """(exact-exchange fraction, is GGA) of a functional name the library knows."""
    hf, gga = C.c_double(), C.c_int()
    if load().oqph_xc_functional_info(functional.encode(), C.byref(hf), C.byref(gga)) != 0:
        raise OqphError(f"unknown functional {functional!r}")
    return hf.value, bool(gga.value)
