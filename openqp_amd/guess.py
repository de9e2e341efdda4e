"""Extended-Hueckel initial guess (host code), restating OpenQP's
guess_huckel (source/modules/guess_huckel.F90:33, source/huckel.F90:31-485,
source/guess.F90:123-300):

1. extended Hueckel in the MINI minimal basis: diagonal = atomic orbital
   energies (core and valence, huckel_lut.F90), off-diagonal
   H_ij = K/2 (H_ii + H_jj) S_ij with K = 1.75, core/core and core/valence
   overlaps first scaled by 0.05; eigenvectors orthonormalised in the MINI
   overlap metric;
2. the first nproj = min(n_mini, ndoc + nact + 5) Hueckel orbitals are mapped
   onto the canonical orthonormal orbitals of the target basis by the
   corresponding-orbital projection (doubly occupied, singly occupied and low
   virtual blocks in turn, each through the SVD of its overlap matrix, then
   paired greedily with the Hueckel orbital of largest overlap);
3. the result is re-orthonormalised in the target overlap metric, the first
   nproj orbitals by QR in order, the rest as their complement.

The overlap between the two bases comes from the one-electron integrals of
the union basis (liboqph's host int1e).  Elements H..Ne.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import basis as B
from .lib import int1e

_HUCKEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "huckel.json")


def _union(b1: B.Basis, b2: B.Basis) -> B.Basis:
    """Shells of b1 then of b2 on the same atoms (one basis for the overlap)."""
    n1 = int(b1.ex.size)
    a1 = b1.nbf
    i32 = lambda v: np.ascontiguousarray(v, np.int32)
    return B.Basis(b1.name + "+" + b2.name, i32(np.concatenate([b1.am, b2.am])),
                   i32(np.concatenate([b1.harmonic, b2.harmonic])), i32(np.concatenate([b1.atom, b2.atom])),
                   i32(np.concatenate([b1.nprim, b2.nprim])), i32(np.concatenate([b1.poff, b2.poff + n1])),
                   i32(np.concatenate([b1.aooff, b2.aooff + a1])), np.concatenate([b1.ex, b2.ex]),
                   np.concatenate([b1.cc, b2.cc]), b1.xyz, b1.zn)


def _canonical(S: np.ndarray, tol: float = 1.0e-8):
    """Canonical orthonormal orbitals Q = U s^-1/2 (Q^T S Q = 1), dropping
    near-dependent combinations."""
    e, U = np.linalg.eigh(S)
    keep = e > tol
    return U[:, keep] / np.sqrt(e[keep])


def _orthonormalise(q: np.ndarray, S: np.ndarray, V: np.ndarray, n: int) -> np.ndarray:
    """orthogonalize_orbitals (huckel.F90:390-440): the first n columns of V
    orthonormalised in order in the S metric, completed by the rest of the
    space q spans; returns (nbf, l0)."""
    W = q.T @ S @ V[:, :n]
    Qf, _ = np.linalg.qr(W, mode="complete")
    return q @ Qf


def huckel_minimal(mini: B.Basis, params: dict):
    """(orbitals (n, n) S-orthonormal, energies ascending) of the extended
    Hueckel operator in the MINI basis (huckel_calc, huckel.F90:146-280)."""
    S, _, _ = int1e(mini)
    n = S.shape[0]
    q = _canonical(S, 0.0)
    H = S.copy()
    core = np.zeros(n, bool)
    lneg = params["lneg_type1"]
    i0 = 0
    for ia, z in enumerate(mini.zn):
        el = params["elements"][str(int(z))]
        en = el["eneg"]
        nbf_atom = int(sum(B.nao_shell(int(l), 0) for l, a in zip(mini.am, mini.atom) if a == ia))
        k = i0
        for i in range(el["ncore"]):
            H[k, k] = en[lneg[i] - 1]
            core[k] = True
            k += 1
        j0 = 0
        for j in range(el["nval"]):
            for i in range(el["ndval"][j]):
                H[k, k] = en[lneg[el["ncore"] + j0 + i] - 1]
                k += 1
            j0 += el["ndval"][j]
        if k != i0 + nbf_atom:
            raise ValueError("Hueckel: MINI basis does not match the orbital table")
        i0 = k
    K, bitsy = params["wh_k"], params["bitsy"]
    for i in range(1, n):
        for j in range(i):
            s = H[j, i] * (bitsy if (core[i] or core[j]) else 1.0)
            H[j, i] = H[i, j] = 0.5 * K * s * (H[i, i] + H[j, j])
    e, V = np.linalg.eigh(H)
    return _orthonormalise(q, S, V, n), e


def _project(va: np.ndarray, rhs: np.ndarray, lo: int, hi: int):
    """One block of corresponding_orbital_projection (guess.F90:172-262):
    rotate va[:, lo:] so its first hi-lo columns overlap the reference
    orbitals lo..hi-1 most, in their order."""
    nrest = va.shape[1] - lo
    nmo = hi - lo
    if nrest <= 0 or nmo <= 0:
        return
    if nmo > nrest:
        raise ValueError("corresponding orbital projection: nmo > nrest")
    D = rhs[lo:hi] @ va[:, lo:]
    _, _, vt = np.linalg.svd(D, full_matrices=False)
    Qf, _ = np.linalg.qr(vt[:nmo].T, mode="complete")
    va[:, lo:] = va[:, lo:] @ Qf
    T = va[:, lo:lo + nmo].T @ rhs[lo:hi].T          # (nmo, nmo)
    unused = np.ones(nmo, bool)
    order = np.zeros(nmo, int)
    for i in range(nmo):
        cand = np.where(unused, np.abs(T[:, i]), -1.0)
        j = int(np.argmax(cand))
        order[i] = j
        unused[j] = False
    va[:, lo:lo + nmo] = va[:, lo:lo + nmo][:, order]


def huckel_guess(bas: B.Basis, ndoc: int, nact: int = 0):
    """Guess orbitals (nbf, l0), S-orthonormal, and approximate orbital
    energies (the Hueckel eigenvalues of the projected orbitals, 0 beyond)."""
    with open(_HUCKEL) as fh:
        params = json.load(fh)
    for z in bas.zn:
        if str(int(z)) not in params["elements"]:
            raise ValueError(f"Hueckel guess: no parameters for Z={int(z)}")
    mini = B.build_basis("mini_huckel", bas.zn.astype(int), bas.xyz)
    vec, heig = huckel_minimal(mini, params)
    U = _union(bas, mini)
    Su, _, _ = int1e(U)
    n = bas.nbf
    S = Su[:n, :n]
    Sco = Su[n:, :n]                                   # (n_mini, nbf)
    nproj = min(vec.shape[1], ndoc + nact + 5)
    q = _canonical(S)
    va = q.copy()
    rhs = vec[:, :nproj].T @ Sco                       # (nproj, nbf)
    _project(va, rhs, 0, ndoc)
    _project(va, rhs, ndoc, ndoc + nact)
    _project(va, rhs, ndoc + nact, nproj)
    C = _orthonormalise(q, S, va, nproj)
    eps = np.zeros(C.shape[1])
    eps[:nproj] = heig[:nproj]
    return C, eps
