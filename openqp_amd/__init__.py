"""openqp_amd -- MI355X (gfx950) direct-SCF two-electron Fock builder (J/K)
and two-electron nuclear gradient, a drop-in for OpenQP's Route-C seam.

Host pieces (basis, synthetic workloads, Python binding) live here; the
compute path is the HIP library openqp_amd/_lib/liboqph.so built from
openqp_amd/csrc by openqp_amd/build.py.
"""
from .basis import Basis, build_basis, pack_tri, unpack_tri  # noqa: F401

__all__ = ["Basis", "build_basis", "pack_tri", "unpack_tri"]
