"""CPU restatement of the semilocal exchange-correlation build (numpy).

Test infrastructure only: imported by tests/ (and the CPU-baseline leg of
bench.py) as the checker of the device Vxc build, never on the product path.

What it restates (the reference's native path behind ``routec_vxc``,
source/scf_addons.F90:1857-1942 -> dmatd_density_blk,
source/dftlib/dft_gridint_energy.F90:408-516):

* input: the packed total density (nfocks = 1, closed shell) or the packed
  alpha and beta densities (nfocks = 2);
* rho(r) = sum_mn D_mn phi_m phi_n, grad rho = 2 sum_mn D_mn phi_m grad phi_n;
* E_xc = sum_g w_g f(rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb);
* V^a_mn = sum_g w_g [df/drho_a phi_m phi_n
  + (2 df/dsigma_aa grad rho_a + df/dsigma_ab grad rho_b) . grad(phi_m phi_n)],
  i.e. dE_xc/dD^a (closed shell: the same matrix for both spins);
* N_elec = sum_g w_g rho.

The functionals are the libxc compositions the reference names
(source/dftlib/libxc.F90:300-460); libxc itself (third party) is absent, so
each component is restated from its published form (Slater exchange; Becke
1988 exchange, beta = 0.0042; VWN correlation, the Ceperley-Alder fit
("VWN5") and the RPA fit; Lee-Yang-Parr correlation in the form of Miehlich
et al. 1989).  Derivatives of the energy density are taken by complex-step
differentiation of the same code (exact to rounding), independently of the
device's forward-mode derivatives.

The grid is this library's own molecular grid (the reference's is built from
its Lebedev tables): per atom a Treutler-Ahlrichs M4 radial grid (Chebyshev
second kind, alpha = 0.6, per-element xi) times a Gauss-Legendre (cos theta)
x uniform (phi) product angular grid (exact for spherical harmonics up to
degree 2 ntheta - 1), with Stratmann-Scuseria-Frisch partition weights
(a = 0.64).  It integrates the same functionals to the grid error only, which
the tests bound against the reference's SCF energies.
"""
from __future__ import annotations

import math

import numpy as np

# Treutler-Ahlrichs radial scaling xi per element (J. Chem. Phys. 102, 346
# (1995), Table I); elements beyond Kr use 1.0
_TA_XI = {1: 0.8, 2: 0.9, 3: 1.8, 4: 1.4, 5: 1.3, 6: 1.1, 7: 0.9, 8: 0.9, 9: 0.9, 10: 0.9,
          11: 1.4, 12: 1.3, 13: 1.3, 14: 1.2, 15: 1.1, 16: 1.0, 17: 1.0, 18: 1.0,
          19: 1.5, 20: 1.4, 21: 1.3, 22: 1.2, 23: 1.2, 24: 1.2, 25: 1.2, 26: 1.2, 27: 1.2,
          28: 1.1, 29: 1.1, 30: 1.1, 31: 1.1, 32: 1.0, 33: 0.9, 34: 0.9, 35: 0.9, 36: 0.9}
SSF_A = 0.64


def ta_xi(z: int) -> float:
    return _TA_XI.get(int(z), 1.0)


def radial_grid(z: int, nrad: int):
    """Treutler-Ahlrichs M4 radial nodes and weights (r^2 dr included)."""
    i = np.arange(1, nrad + 1)
    th = i * math.pi / (nrad + 1)
    x = np.cos(th)
    xi, al = ta_xi(z), 0.6
    c = xi / math.log(2.0)
    lg = np.log(2.0 / (1.0 - x))
    r = c * (1.0 + x) ** al * lg
    drdx = c * (al * (1.0 + x) ** (al - 1.0) * lg + (1.0 + x) ** al / (1.0 - x))
    w = math.pi / (nrad + 1) * np.sin(th) * drdx * r * r
    return r, w


def angular_grid(ntheta: int):
    """Unit vectors and weights (sum 4 pi) of the product angular grid."""
    t, u = np.polynomial.legendre.leggauss(ntheta)
    nphi = 2 * ntheta
    phi = (np.arange(nphi) + 0.5) * (2.0 * math.pi / nphi)
    st = np.sqrt(1.0 - t * t)
    v = np.stack([np.outer(st, np.cos(phi)), np.outer(st, np.sin(phi)), np.outer(t, np.ones(nphi))], -1)
    w = np.outer(u, np.full(nphi, 2.0 * math.pi / nphi))
    return v.reshape(-1, 3), w.reshape(-1)


def ssf_weights(pts: np.ndarray, owner: int, xyz: np.ndarray) -> np.ndarray:
    """Stratmann-Scuseria-Frisch partition weight of atom ``owner``."""
    natm = xyz.shape[0]
    if natm == 1:
        return np.ones(len(pts))
    d = np.linalg.norm(pts[:, None, :] - xyz[None, :, :], axis=2)  # (n, natm)
    P = np.ones((len(pts), natm))
    for b in range(natm):
        for c in range(natm):
            if b == c:
                continue
            R = np.linalg.norm(xyz[b] - xyz[c])
            mu = (d[:, b] - d[:, c]) / R
            x = np.clip(mu / SSF_A, -1.0, 1.0)
            g = (35 * x - 35 * x ** 3 + 21 * x ** 5 - 5 * x ** 7) / 16.0
            s = 0.5 * (1.0 - g)
            s = np.where(mu <= -SSF_A, 1.0, np.where(mu >= SSF_A, 0.0, s))
            P[:, b] *= s
    tot = P.sum(1)
    return np.where(tot > 0, P[:, owner] / np.where(tot > 0, tot, 1.0), 0.0)


def molecular_grid(zn, xyz, nrad=96, ntheta=15, wcut=1e-20):
    """Points (n, 3), weights (n,) and owning atom (n,), atom-major, radial-major."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    av, aw = angular_grid(ntheta)
    P, W, A = [], [], []
    for a in range(xyz.shape[0]):
        r, wr = radial_grid(int(round(zn[a])), nrad)
        pts = (xyz[a][None, None, :] + r[:, None, None] * av[None, :, :]).reshape(-1, 3)
        w = (wr[:, None] * aw[None, :]).reshape(-1) * ssf_weights(pts, a, xyz)
        keep = w > wcut
        P.append(pts[keep])
        W.append(w[keep])
        A.append(np.full(int(keep.sum()), a, dtype=np.int32))
    return np.concatenate(P), np.concatenate(W), np.concatenate(A)


# ---------------------------------------------------------------------------
# basis functions on the grid

def _dfact(n: int) -> float:
    return float(np.prod(np.arange(n, 0, -2))) if n > 0 else 1.0


def ao_values(bas, pts, oracle):
    """phi (n, nbf) and grad phi (3, n, nbf) in OpenQP's AO representation
    (component order, shells_pnrm2 norms, pure projection for l >= 2)."""
    n = len(pts)
    nbf = bas.nbf
    phi = np.zeros((n, nbf))
    dphi = np.zeros((3, n, nbf))
    for s in range(bas.nsh):
        l = int(bas.am[s])
        A = bas.xyz[int(bas.atom[s])]
        p0, npr = int(bas.poff[s]), int(bas.nprim[s])
        ex = bas.ex[p0:p0 + npr]
        cc = bas.cc[p0:p0 + npr]
        d = pts - A
        r2 = np.einsum("ij,ij->i", d, d)
        e = np.exp(-np.outer(r2, ex))             # (n, nprim)
        R = e @ cc
        dR = -2.0 * (e @ (cc * ex))               # (1/r) dR/dr
        comps = oracle.cart_order(l)
        nc = len(comps)
        cart = np.zeros((n, nc))
        dcart = np.zeros((3, n, nc))
        for q, (a, b, c) in enumerate(comps):
            nrm = math.sqrt(_dfact(2 * l - 1) / (_dfact(2 * a - 1) * _dfact(2 * b - 1) * _dfact(2 * c - 1)))
            mono = d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c
            cart[:, q] = nrm * mono * R
            for k, (pw) in enumerate((a, b, c)):
                dm = np.zeros(n)
                if pw > 0:
                    ex3 = [a, b, c]
                    ex3[k] -= 1
                    dm = pw * d[:, 0] ** ex3[0] * d[:, 1] ** ex3[1] * d[:, 2] ** ex3[2]
                dcart[k, :, q] = nrm * (dm * R + mono * dR * d[:, k])
        o = int(bas.aooff[s])
        if int(bas.harmonic[s]) and l >= 2:
            T = oracle.c2s(l)                       # (ncart, 2l+1)
            no = 2 * l + 1
            phi[:, o:o + no] = cart @ T
            dphi[:, :, o:o + no] = dcart @ T
        else:
            phi[:, o:o + nc] = cart
            dphi[:, :, o:o + nc] = dcart
    return phi, dphi


# ---------------------------------------------------------------------------
# functionals: energy density f(ra, rb, saa, sab, sbb) per unit volume

_VWN5 = ((0.0310907, 3.72744, 12.9352, -0.10498),
         (0.01554535, 7.06042, 18.0578, -0.32500),
         (-1.0 / (6.0 * math.pi ** 2), 1.13107, 13.0045, -0.0047584))
_VWNRPA = ((0.0310907, 13.0720, 42.7198, -0.409286),
           (0.01554535, 20.1231, 101.578, -0.743294))
LYP_A, LYP_B, LYP_C, LYP_D = 0.04918, 0.132, 0.2533, 0.349
B88_BETA = 0.0042
CX = 0.75 * (6.0 / math.pi) ** (1.0 / 3.0)


def lda_x(ra, rb, saa, sab, sbb):
    return -CX * (ra ** (4.0 / 3.0) + rb ** (4.0 / 3.0))


def b88(ra, rb, saa, sab, sbb):
    out = 0.0
    for r, s in ((ra, saa), (rb, sbb)):
        r43 = r ** (4.0 / 3.0)
        x = np.sqrt(s) / r43
        out = out - CX * r43 - B88_BETA * r43 * x * x / (1.0 + 6.0 * B88_BETA * x * np.arcsinh(x))
    return out


def _vwn_aux(p, rs):
    A, b, c, x0 = p
    x = np.sqrt(rs)
    X = x * x + b * x + c
    X0 = x0 * x0 + b * x0 + c
    Q = math.sqrt(4.0 * c - b * b)
    at = np.arctan(Q / (2.0 * x + b))
    return A * (np.log(x * x / X) + 2.0 * b / Q * at
                - b * x0 / X0 * (np.log((x - x0) ** 2 / X) + 2.0 * (b + 2.0 * x0) / Q * at))


def _fz(z):
    return ((1.0 + z) ** (4.0 / 3.0) + (1.0 - z) ** (4.0 / 3.0) - 2.0) / (2.0 ** (4.0 / 3.0) - 2.0)


FPP0 = 4.0 / (9.0 * (2.0 ** (1.0 / 3.0) - 1.0))


def vwn5(ra, rb, saa, sab, sbb):
    r = ra + rb
    rs = (3.0 / (4.0 * math.pi * r)) ** (1.0 / 3.0)
    z = (ra - rb) / r
    eP, eF, ac = (_vwn_aux(p, rs) for p in _VWN5)
    fz = _fz(z)
    z4 = z ** 4
    return r * (eP + ac * fz / FPP0 * (1.0 - z4) + (eF - eP) * fz * z4)


def vwn_rpa(ra, rb, saa, sab, sbb):
    r = ra + rb
    rs = (3.0 / (4.0 * math.pi * r)) ** (1.0 / 3.0)
    z = (ra - rb) / r
    eP, eF = (_vwn_aux(p, rs) for p in _VWNRPA)
    fz = _fz(z)
    return r * (eP * (1.0 - fz) + eF * fz)


def lyp(ra, rb, saa, sab, sbb):
    a, b, c, d = LYP_A, LYP_B, LYP_C, LYP_D
    r = ra + rb
    rm = r ** (-1.0 / 3.0)
    den = 1.0 + d * rm
    om = np.exp(-c * rm) / den * r ** (-11.0 / 3.0)
    dl = c * rm + d * rm / den
    cf = 0.3 * (3.0 * math.pi ** 2) ** (2.0 / 3.0)
    s = saa + 2.0 * sab + sbb
    t = ra * rb * (2.0 ** (11.0 / 3.0) * cf * (ra ** (8.0 / 3.0) + rb ** (8.0 / 3.0))
                   + (47.0 / 18.0 - 7.0 / 18.0 * dl) * s - (2.5 - dl / 18.0) * (saa + sbb)
                   - (dl - 11.0) / 9.0 * (ra / r * saa + rb / r * sbb))
    t = t - 2.0 / 3.0 * r * r * s + (2.0 / 3.0 * r * r - ra * ra) * sbb + (2.0 / 3.0 * r * r - rb * rb) * saa
    return -4.0 * a / den * ra * rb / r - a * b * om * t


COMPONENTS = {"lda_x": (lda_x, False), "b88": (b88, True), "vwn5": (vwn5, False),
              "vwn_rpa": (vwn_rpa, False), "lyp": (lyp, True)}

# name -> ([(component, weight)], exact-exchange fraction); the libxc
# compositions selected in source/dftlib/libxc.F90:300-460
FUNCTIONALS = {
    "slater": ([("lda_x", 1.0)], 0.0),
    "svwn5": ([("lda_x", 1.0), ("vwn5", 1.0)], 0.0),
    "blyp": ([("b88", 1.0), ("lyp", 1.0)], 0.0),
    "b3lyp5": ([("lda_x", 0.08), ("b88", 0.72), ("vwn5", 0.19), ("lyp", 0.81)], 0.20),
    "b3lyp": ([("lda_x", 0.08), ("b88", 0.72), ("vwn_rpa", 0.19), ("lyp", 0.81)], 0.20),
    "bhhlyp": ([("b88", 0.5), ("lyp", 1.0)], 0.5),
}


def canonical_name(name: str) -> str:
    n = name.lower()
    return {"b3lypv5": "b3lyp5", "b3lypv1r": "b3lyp", "bhandhlyp": "bhhlyp"}.get(n, n)


def energy_density(name, ra, rb, saa, sab, sbb):
    comps, _ = FUNCTIONALS[canonical_name(name)]
    f = 0.0
    for c, w in comps:
        f = f + w * COMPONENTS[c][0](ra, rb, saa, sab, sbb)
    return f


def is_gga(name) -> bool:
    return any(COMPONENTS[c][1] for c, _ in FUNCTIONALS[canonical_name(name)][0])


def hf_exchange(name) -> float:
    return FUNCTIONALS[canonical_name(name)][1]


def derivatives(fun, args):
    """Value and complex-step partials of fun(*args) with respect to each arg."""
    args = [np.asarray(a, dtype=np.float64) for a in args]
    val = np.real(fun(*args))
    out = []
    for k, a in enumerate(args):
        h = 1e-20 * np.maximum(np.abs(a), 1e-200)
        z = [x.astype(np.complex128) for x in args]
        z[k] = z[k] + 1j * h
        out.append(np.imag(fun(*z)) / h)
    return val, out


RHO_CUT = 1e-10   # points whose total density is below this contribute nothing


def vxc(bas, dens, name, grid, oracle, nfocks=1, ao=None):
    """Vxc (nfocks full matrices), E_xc and N_elec on ``grid`` = (pts, w).

    nfocks = 1: dens = full total density; nfocks = 2: (D_alpha, D_beta).
    ao: the basis functions on the grid from ao_values, when already known."""
    pts, w = grid
    phi, dphi = ao if ao is not None else ao_values(bas, pts, oracle)
    gga = is_gga(name)
    Ds = [dens] if nfocks == 1 else list(dens)
    rho, grad = [], []
    for D in Ds:
        X = phi @ D
        rho.append(np.einsum("gm,gm->g", X, phi))
        grad.append(2.0 * np.einsum("gm,kgm->kg", X, dphi))
    if nfocks == 1:
        ra = rb = 0.5 * rho[0]
        ga = gb = 0.5 * grad[0]
        rtot = rho[0]
    else:
        ra, rb = rho
        ga, gb = grad
        rtot = ra + rb
    ok = rtot >= RHO_CUT
    ra, rb = np.where(ok, ra, 1.0), np.where(ok, rb, 1.0)
    saa = np.einsum("kg,kg->g", ga, ga)
    sab = np.einsum("kg,kg->g", ga, gb)
    sbb = np.einsum("kg,kg->g", gb, gb)
    ra_, rb_ = np.maximum(ra, 1e-300), np.maximum(rb, 1e-300)
    f, (fra, frb, fsaa, fsab, fsbb) = derivatives(lambda *a: energy_density(name, *a), [ra_, rb_, saa, sab, sbb])
    wz = np.where(ok, w, 0.0)
    exc = float(np.sum(wz * f))
    nel = float(np.sum(w * rtot))
    V = []
    spins = [(fra, fsaa, fsab, ga, gb)] if nfocks == 1 else [(fra, fsaa, fsab, ga, gb), (frb, fsbb, fsab, gb, ga)]
    for fr, fss, fsx, gs, go in spins:
        M = phi.T @ (phi * (wz * fr)[:, None])
        if gga:
            vec = 2.0 * fss[None, :] * gs + fsx[None, :] * go          # (3, n)
            Y = np.einsum("kg,kgm->gm", vec * wz[None, :], dphi)
            T = phi.T @ Y
            M = M + T + T.T
        V.append(M)
    return V, exc, nel
