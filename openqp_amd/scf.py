"""Host RHF SCF loop around the device Fock builder.

Restates the semantics of OpenQP's scf_driver for closed shells
(source/scf.F90:46-1364, source/scf_addons.F90:1972-2145):

* Fock: F = H + G[D], G from the device build (packed, fock_jk semantics);
* energy: E = 0.5 (Tr D F + Tr D H) + E_nuc (scf_addons.F90:2059-2065);
* incremental builds: F = F_old + G[D - D_old]; the history is reset every
  10th iteration and whenever the DIIS error is below 1e-4
  (scf.F90:768-773);
* convergence: DIIS error max|X^T (F D S - S D F) X| below ``conv`` and the
  energy change below ``econv``;
* Pulay DIIS on the orthonormal-basis commutator; the B matrix is scaled by
  its largest diagonal entry and the oldest vectors are dropped while it is
  ill-conditioned (reciprocal condition number below 1e-12), so the
  extrapolation coefficients stay meaningful near convergence.

Kohn-Sham (``functional=``, closed shell): F = H + G_se[D] + Vxc[D] with the
exact-exchange scale se of the functional and Vxc from the builder's XC
build (the routec_vxc work), E = 0.5 (Tr D (H + G) + Tr D H) + E_xc + E_nuc,
as calc_jk_xc assembles them (source/scf_addons.F90:1972-2145).

Initial guess: OpenQP's extended-Hueckel guess (``guess="huckel"``,
source/modules/guess_huckel.F90:33; openqp_amd/guess.py) or the core
Hamiltonian (``guess="core"``).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import basis as B
from .lib import FockBuilder, int1e


@dataclass
class SCFResult:
    energy: float
    converged: bool
    iterations: int
    density: np.ndarray          # full total density
    fock: np.ndarray             # full Fock matrix
    orbitals: np.ndarray
    orbital_energies: np.ndarray
    fock_ms: list = field(default_factory=list)
    history: list = field(default_factory=list)
    exc: float = 0.0             # E_xc (Kohn-Sham)
    nelec_grid: float = 0.0      # the XC grid's electron count (Kohn-Sham)


class RHF:
    def __init__(self, bas: B.Basis, builder: FockBuilder | None = None, nocc: int | None = None,
                 conv: float = 1e-6, econv: float = 1e-8, max_iter: int = 100, incremental: bool = True,
                 diis_size: int = 8, guess: str = "huckel", functional: str | None = None,
                 grid: tuple = (96, 15)):
        self.bas = bas
        self.builder = builder or FockBuilder(bas)
        self.functional = functional
        self.hfex = 1.0
        if functional is not None:
            from .lib import xc_functional_info
            self.hfex = xc_functional_info(functional)[0]
            self.builder.set_xc(functional, *grid)
        self.nocc = nocc if nocc is not None else int(round(bas.zn.sum())) // 2
        self.conv, self.econv, self.max_iter = conv, econv, max_iter
        self.incremental = incremental
        self.diis_size = diis_size
        if guess not in ("huckel", "core"):
            raise ValueError(f"unknown guess {guess!r} (huckel, core)")
        self.guess = guess
        S, T, V = int1e(bas)
        self.S, self.H = S, T + V
        self.enuc = B.nuclear_repulsion(bas.zn, bas.xyz)
        e, U = np.linalg.eigh(S)
        keep = e > 1e-8
        self.X = U[:, keep] / np.sqrt(e[keep])

    def _density(self, F):
        Fo = self.X.T @ F @ self.X
        e, Co = np.linalg.eigh(Fo)
        C = self.X @ Co
        Cocc = C[:, :self.nocc]
        return 2.0 * Cocc @ Cocc.T, C, e

    def initial_density(self) -> np.ndarray:
        if self.guess == "huckel":
            from .guess import huckel_guess
            C, _ = huckel_guess(self.bas, self.nocc)
            Cocc = C[:, :self.nocc]
            return 2.0 * Cocc @ Cocc.T
        return self._density(self.H)[0]

    @staticmethod
    def diis_coefficients(errs) -> np.ndarray:
        """Pulay coefficients for the stored error vectors: B scaled by its
        largest diagonal entry; the oldest vectors are dropped while B is
        ill-conditioned (rcond < 1e-12).  Returns one coefficient per stored
        vector (zeros for the dropped ones)."""
        n = len(errs)
        G = np.array([[float(np.sum(errs[i] * errs[j])) for j in range(n)] for i in range(n)])
        for k in range(n - 1):
            m = n - k
            Gk = G[k:, k:]
            scale = float(np.max(np.diag(Gk)))
            Bm = -np.ones((m + 1, m + 1))
            Bm[m, m] = 0.0
            Bm[:m, :m] = Gk / scale
            if 1.0 / np.linalg.cond(Bm) < 1.0e-12:
                continue
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.zeros(n)
            c[k:] = np.linalg.solve(Bm, rhs)[:m]
            return c
        c = np.zeros(n)
        c[-1] = 1.0
        return c

    def run(self, D0: np.ndarray | None = None) -> SCFResult:
        nbf = self.bas.nbf
        D = self.initial_density() if D0 is None else D0
        dp = B.pack_tri(D)
        d_old = np.zeros_like(dp)
        f_old = np.zeros_like(dp)
        errs, focks = [], []
        e_prev, diis_err = 0.0, 1.0
        res = SCFResult(0.0, False, 0, D, self.H, None, None)
        for it in range(1, self.max_iter + 1):
            t0 = time.perf_counter()
            if self.incremental:
                if it > 1 and (it % 10 == 0 or diis_err < 1.0e-4):
                    d_old[:] = 0.0
                    f_old[:] = 0.0
                g = self.builder.fock_jk_inc(dp, d_old, f_old, se=self.hfex)
            else:
                g = self.builder.fock_jk(dp, se=self.hfex)
            G = B.unpack_tri(g, nbf)
            E = 0.5 * (np.sum(D * G) + 2.0 * np.sum(D * self.H)) + self.enuc
            F = self.H + G
            if self.functional is not None:
                vx, exc, nel = self.builder.vxc(dp, 1)
                F = F + B.unpack_tri(vx[0], nbf)
                E += exc
                res.exc, res.nelec_grid = exc, nel
            res.fock_ms.append((time.perf_counter() - t0) * 1e3)
            FDS = F @ D @ self.S  # S D F is its transpose (F, D, S symmetric)
            err = self.X.T @ (FDS - FDS.T) @ self.X
            diis_err = float(np.abs(err).max())
            res.history.append((it, E, diis_err))
            if it > 1 and diis_err < self.conv and abs(E - e_prev) < self.econv:
                res.energy, res.converged, res.iterations = E, True, it
                res.density, res.fock = D, F
                _, res.orbitals, res.orbital_energies = self._density(F)
                return res
            e_prev = E
            focks.append(F)
            errs.append(err)
            if len(focks) > self.diis_size:
                focks.pop(0)
                errs.pop(0)
            if len(focks) > 1:
                c = self.diis_coefficients(errs)
                Fx = sum(ci * Fi for ci, Fi in zip(c, focks) if ci != 0.0)
            else:
                Fx = F
            D, C, eps = self._density(Fx)
            dp = B.pack_tri(D)
        res.energy, res.iterations = E, self.max_iter
        res.density, res.fock = D, F
        return res


class UHF(RHF):
    """Unrestricted host loop (UHF, or UKS with ``functional=``): the alpha and
    beta Fock matrices from one nfocks = 2 build (int2_urohf_data_t semantics,
    source/integrals/int2.F90:1488-1578) plus the alpha/beta Vxc,
    E = 0.5 sum_s (Tr D_s (H + G_s) + Tr D_s H) + E_xc + E_nuc, Pulay DIIS on
    the stacked alpha and beta commutators.  Full builds every iteration.

    Starts from orbitals C0 when given (alpha and beta share them), else from
    the guess; ``nalpha``/``nbeta`` default to the multiplicity's occupations."""

    def __init__(self, bas: B.Basis, builder: FockBuilder | None = None, mult: int = 1,
                 nalpha: int | None = None, nbeta: int | None = None, **kw):
        nel = int(round(bas.zn.sum()))
        na = nalpha if nalpha is not None else (nel + mult - 1) // 2
        nb = nbeta if nbeta is not None else nel - na
        super().__init__(bas, builder, nocc=nb, **kw)
        self.nalpha, self.nbeta = na, nb

    def _orbitals(self, F):
        e, Co = np.linalg.eigh(self.X.T @ F @ self.X)
        return self.X @ Co, e

    def run(self, C0: np.ndarray | None = None) -> SCFResult:
        nbf = self.bas.nbf
        if C0 is None:
            if self.guess == "huckel":
                from .guess import huckel_guess
                C0, _ = huckel_guess(self.bas, self.nalpha)
            else:
                C0, _ = self._orbitals(self.H)
        Ds = [C0[:, :n] @ C0[:, :n].T for n in (self.nalpha, self.nbeta)]
        errs, focks = [], []
        e_prev = 0.0
        res = SCFResult(0.0, False, 0, Ds, self.H, None, None)
        for it in range(1, self.max_iter + 1):
            t0 = time.perf_counter()
            dps = [B.pack_tri(D) for D in Ds]
            ga, gb = self.builder.fock_jk_uhf(dps[0], dps[1], se=self.hfex)
            G = [B.unpack_tri(ga, nbf), B.unpack_tri(gb, nbf)]
            E = sum(0.5 * (np.sum(D * g) + 2.0 * np.sum(D * self.H)) for D, g in zip(Ds, G)) + self.enuc
            Fs = [self.H + g for g in G]
            if self.functional is not None:
                vx, exc, nel = self.builder.vxc(np.concatenate(dps), 2)
                Fs = [F + B.unpack_tri(v, nbf) for F, v in zip(Fs, vx)]
                E += exc
                res.exc, res.nelec_grid = exc, nel
            res.fock_ms.append((time.perf_counter() - t0) * 1e3)
            err = []
            for F, D in zip(Fs, Ds):
                FDS = F @ D @ self.S
                err.append(self.X.T @ (FDS - FDS.T) @ self.X)
            err = np.stack(err)
            diis_err = float(np.abs(err).max())
            res.history.append((it, E, diis_err))
            if it > 1 and diis_err < self.conv and abs(E - e_prev) < self.econv:
                res.energy, res.converged, res.iterations = E, True, it
                res.density, res.fock = Ds, Fs
                res.orbitals, res.orbital_energies = zip(*[self._orbitals(F) for F in Fs])
                return res
            e_prev = E
            focks.append(Fs)
            errs.append(err)
            if len(focks) > self.diis_size:
                focks.pop(0)
                errs.pop(0)
            if len(focks) > 1:
                c = self.diis_coefficients(errs)
                Fx = [sum(ci * f[s] for ci, f in zip(c, focks) if ci != 0.0) for s in range(2)]
            else:
                Fx = Fs
            Ds = []
            for F, n in zip(Fx, (self.nalpha, self.nbeta)):
                C, _ = self._orbitals(F)
                Ds.append(C[:, :n] @ C[:, :n].T)
        res.energy, res.iterations = E, self.max_iter
        res.density, res.fock = Ds, Fs
        return res
