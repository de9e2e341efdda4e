"""Shared test helpers: workloads, densities, header parsing."""
import os
import re

import numpy as np

from openqp_amd import basis as B, molecules as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def workload(name, seed=0):
    bld, bname, nocc = M.WORKLOADS[name]
    Z, xyz = bld(seed)
    return B.build_basis(bname, Z, xyz), nocc


def small_system(basis_name, nw=1, seed=0):
    if nw == 1:
        Z, xyz = M.h2o_kat()
    else:
        Z, xyz = M.water_grid(nw, 1, 1, seed=seed)
    return B.build_basis(basis_name, Z, xyz)


def density_for(oracle, bas, nocc, seed=0):
    S, _, _ = oracle.int1e(bas)
    D = M.random_density(S, nocc, seed=seed)
    return D, B.pack_tri(D)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "oqph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|long long|const char \*)\s*\**\s*(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


class DeviceBuffer:
    """Device memory through the HIP runtime liboqph itself links (ctypes on
    /opt/rocm's libamdhip64), so a test can hand device pointers to the C ABI
    without a second HIP runtime in the process."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, n_doubles: int):
        import ctypes as C
        self.n = int(n_doubles)
        self.ptr = C.c_void_p()
        rc = self.hip().hipMalloc(C.byref(self.ptr), max(self.n, 1) * 8)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        rc = self.hip().hipMemcpy(self.ptr, a.ctypes.data, self.n * 8, 1)  # host to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")

    def download(self):
        out = np.empty(self.n)
        self.hip().hipDeviceSynchronize()
        rc = self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, 2)  # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
