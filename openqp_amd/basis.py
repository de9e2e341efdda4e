"""Basis-set representation for the Fock builder (host side).

Mirrors the fields of OpenQP's ``basis_set`` type (source/basis_tools.F90:32-84)
that the two-electron path reads: per-shell angular momentum, harmonic flag,
atom, primitive count and offset, AO offset; per-primitive exponents and
coefficients.  Conventions (source/constants.F90:34-60, 110+;
source/basis_tools.F90:277-303):

* shells follow atom order, then the order of the basis file;
* Cartesian components in OpenQP order (d = xx,yy,zz,xy,xz,yz, ...),
  pure components m = -l..l (source/integrals/cart2sph.F90:1-19);
  s and p shells are never pure;
* each primitive coefficient carries the x^l primitive normalisation
  (``normalize_primitives``); contracted shells are not renormalised --
  pyoqp builds the basis through basis_api.F90:316-317, which calls
  ``normalize_primitives`` only.  Non-axial Cartesian components pick up
  their ``shells_pnrm2`` factor inside the integral engines.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")

# (2l-1)!!/2^l * pi^(3/2) (source/basis_tools.F90:612-623, gauss_norm)
_NORMS = [1.0, 0.5, 0.75, 1.875, 6.5625, 29.53125, 162.421875]


def gauss_norm(e: float, l: int) -> float:
    """Self-overlap integral of the x^l primitive with exponent e/2."""
    return math.pi * math.sqrt(math.pi) * _NORMS[l] / (e * math.sqrt(e) * e ** l)


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def nao_shell(l: int, harmonic: int) -> int:
    return 2 * l + 1 if (harmonic and l >= 2) else ncart(l)


def available_bases():
    return sorted(f[:-5] for f in os.listdir(_DATA) if f.endswith(".json"))


def load_basis_data(name: str) -> dict:
    key = name.lower().replace("*", "(d)")
    path = os.path.join(_DATA, key + ".json")
    if not os.path.exists(path):
        raise ValueError(f"basis '{name}' not shipped (have {available_bases()})")
    with open(path) as fh:
        return json.load(fh)


@dataclass
class Basis:
    """Flat arrays describing a molecular basis (0-based offsets)."""

    name: str
    am: np.ndarray        # int32 [nsh]
    harmonic: np.ndarray  # int32 [nsh]
    atom: np.ndarray      # int32 [nsh]
    nprim: np.ndarray     # int32 [nsh]
    poff: np.ndarray      # int32 [nsh]
    aooff: np.ndarray     # int32 [nsh]
    ex: np.ndarray        # float64 [nprim_total]
    cc: np.ndarray        # float64 [nprim_total], primitive-normalised
    xyz: np.ndarray       # float64 [natm, 3] bohr
    zn: np.ndarray        # float64 [natm]

    @property
    def nsh(self) -> int:
        return int(self.am.size)

    @property
    def natm(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nbf(self) -> int:
        return int(sum(nao_shell(int(l), int(h)) for l, h in zip(self.am, self.harmonic)))

    @property
    def ntri(self) -> int:
        n = self.nbf
        return n * (n + 1) // 2

    @property
    def lmax(self) -> int:
        return int(self.am.max())

    def naos(self) -> np.ndarray:
        return np.array([nao_shell(int(l), int(h)) for l, h in zip(self.am, self.harmonic)], dtype=np.int32)

    def with_coords(self, xyz: np.ndarray) -> "Basis":
        """Same basis on moved nuclei (bohr)."""
        return Basis(self.name, self.am, self.harmonic, self.atom, self.nprim, self.poff, self.aooff,
                     self.ex, self.cc, np.ascontiguousarray(xyz, dtype=np.float64), self.zn)

    def carray(self):
        """Tuple of contiguous arrays in the order the C entry points take them."""
        return (self.nsh, self.nbf, self.natm,
                np.ascontiguousarray(self.am, np.int32), np.ascontiguousarray(self.harmonic, np.int32),
                np.ascontiguousarray(self.atom, np.int32), np.ascontiguousarray(self.nprim, np.int32),
                np.ascontiguousarray(self.poff, np.int32), np.ascontiguousarray(self.aooff, np.int32),
                np.ascontiguousarray(self.ex, np.float64), np.ascontiguousarray(self.cc, np.float64),
                np.ascontiguousarray(self.xyz, np.float64))


def build_basis(name: str, Z, xyz_bohr, harmonic: int | None = None) -> Basis:
    """Assemble the molecular basis ``name`` on atoms ``Z`` at ``xyz_bohr``."""
    data = load_basis_data(name)
    harm = data["harmonic"] if harmonic is None else int(harmonic)
    am, hf, at, npr, po, ao, ex, cc = [], [], [], [], [], [], [], []
    p = a = 0
    for ia, z in enumerate(Z):
        shells = data["elements"].get(str(int(z)))
        if shells is None:
            raise ValueError(f"basis {name} has no element Z={z}")
        for sh in shells:
            l = int(sh["l"])
            h = 1 if (harm and l >= 2) else 0
            e = np.asarray(sh["exps"], dtype=np.float64)
            c = np.asarray(sh["coefs"], dtype=np.float64)
            c = c / np.sqrt(np.array([gauss_norm(2.0 * x, l) for x in e]))
            am.append(l); hf.append(h); at.append(ia); npr.append(e.size); po.append(p); ao.append(a)
            ex.extend(e.tolist()); cc.extend(c.tolist())
            p += e.size
            a += nao_shell(l, h)
    i32 = lambda v: np.asarray(v, dtype=np.int32)
    return Basis(name, i32(am), i32(hf), i32(at), i32(npr), i32(po), i32(ao),
                 np.asarray(ex, np.float64), np.asarray(cc, np.float64),
                 np.ascontiguousarray(np.asarray(xyz_bohr, np.float64).reshape(-1, 3)),
                 np.asarray(Z, np.float64))


def pack_tri(m: np.ndarray) -> np.ndarray:
    """Row-packed lower triangle, ij = i(i-1)/2 + j (1-based), int2.F90:1436-1447."""
    n = m.shape[0]
    il = np.tril_indices(n)
    # tril_indices is row-major over (i, j<=i): exactly the packed order
    return np.ascontiguousarray(m[il])


def unpack_tri(p: np.ndarray, n: int) -> np.ndarray:
    m = np.zeros((n, n), dtype=p.dtype)
    il = np.tril_indices(n)
    m[il] = p
    m.T[il] = p
    return m


def nuclear_repulsion(Z, xyz) -> float:
    Z = np.asarray(Z, np.float64)
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    e = 0.0
    for i in range(len(Z)):
        for j in range(i):
            e += Z[i] * Z[j] / np.linalg.norm(xyz[i] - xyz[j])
    return float(e)
