"""Seeded synthetic molecules and densities for the Fock-build workloads.

Geometry conventions (all lengths returned in bohr):

* water monomer: O at the origin, H1 at (0.9572 A, 0, 0), H2 at 104.52 deg in
  the xy plane; each molecule gets a uniform random rotation drawn from the
  seed and its O sits on a cubic grid node (3.0 A spacing);
* linear alkane C_nH_{2n+2}: all-trans zig-zag, C-C 1.54 A, C-H 1.09 A,
  tetrahedral angles, plus seeded Gaussian noise (sigma 0.05 A) on every
  coordinate;
* bohr = 0.529177210903 A (source/physical_constants.F90:12).

Densities: C_occ = seeded standard-normal (nbf, nocc), orthonormalised in
the overlap metric (C <- C L^-T with L the Cholesky factor of C^T S C), and
D = 2 C_occ C_occ^T (the total RHF density held in OpenQP's DM_A).
"""
from __future__ import annotations

import numpy as np

BOHR = 0.529177210903


def _random_rotations(n: int, seed: int) -> np.ndarray:
    from scipy.spatial.transform import Rotation
    return Rotation.random(n, random_state=seed).as_matrix()


def water_monomer_angstrom() -> np.ndarray:
    r, th = 0.9572, np.deg2rad(104.52)
    return np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], [r * np.cos(th), r * np.sin(th), 0.0]])


def water_grid(nx: int, ny: int, nz: int, spacing: float = 3.0, seed: int = 0):
    """(Z, xyz_bohr) for nx*ny*nz waters on a cubic grid with random orientations."""
    mono = water_monomer_angstrom()
    n = nx * ny * nz
    rots = _random_rotations(n, seed)
    Z, xyz = [], []
    m = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                node = np.array([ix, iy, iz], dtype=np.float64) * spacing
                pts = mono @ rots[m].T + node
                Z.extend([8, 1, 1])
                xyz.extend(pts.tolist())
                m += 1
    return np.array(Z, dtype=np.int32), np.array(xyz) / BOHR


def alkane(nc: int = 20, seed: int = 0, noise: float = 0.05):
    """(Z, xyz_bohr) for the all-trans linear alkane C_nc H_(2nc+2) with seeded noise."""
    cc, ch = 1.54, 1.09
    half = np.deg2rad(109.4712206) / 2.0
    dx = cc * np.sin(half)
    dz = cc * np.cos(half)
    carbons = np.array([[i * dx, 0.0, (dz / 2.0) * (1 if i % 2 else -1)] for i in range(nc)])
    Z = [6] * nc
    xyz = [c.tolist() for c in carbons]
    # tetrahedral H placement: two H per carbon off the zig-zag plane, plus
    # one extra H on each terminal carbon continuing the chain direction.
    for i, c in enumerate(carbons):
        up = 1.0 if i % 2 == 0 else -1.0  # outward side in the zig-zag plane (z)
        # the two C-H bonds bisect the outward direction and straddle the plane
        tet = np.deg2rad(109.4712206)
        out = np.array([0.0, 0.0, -up])
        v1 = np.cos(tet / 2) * out + np.sin(tet / 2) * np.array([0.0, 1.0, 0.0])
        v2 = np.cos(tet / 2) * out - np.sin(tet / 2) * np.array([0.0, 1.0, 0.0])
        for v in (v1, v2):
            Z.append(1)
            xyz.append((c + ch * v).tolist())
        if i == 0 or i == nc - 1:
            nb = carbons[1] if i == 0 else carbons[nc - 2]
            d = c - nb
            d /= np.linalg.norm(d)
            # continue the zig-zag: reflect the neighbour bond through the outward axis
            v = np.array([d[0], 0.0, -d[2]])
            v /= np.linalg.norm(v)
            Z.append(1)
            xyz.append((c + ch * v).tolist())
    xyz = np.array(xyz)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.normal(0.0, noise, size=xyz.shape)
    return np.array(Z, dtype=np.int32), xyz / BOHR


def h2o_kat():
    """H2O geometry of the OpenQP HF known-answer decks (examples/HF/H2O_RHF-HF_*.inp), bohr."""
    ang = np.array([[0.000000000, 0.000000000, -0.041061554],
                    [-0.533194329, 0.533194329, -0.614469223],
                    [0.533194329, -0.533194329, -0.614469223]])
    return np.array([8, 1, 1], dtype=np.int32), ang / BOHR


def random_density(S: np.ndarray, nocc: int, seed: int = 0) -> np.ndarray:
    """Full symmetric D = 2 C C^T with C S-orthonormal, seeded."""
    rng = np.random.default_rng(seed)
    n = S.shape[0]
    C = rng.standard_normal((n, nocc))
    M = C.T @ S @ C
    L = np.linalg.cholesky(M)
    C = C @ np.linalg.inv(L).T
    return 2.0 * C @ C.T


WORKLOADS = {
    # name: (builder, basis, nocc)
    "h2o8_631g": (lambda seed=0: water_grid(2, 2, 2, seed=seed), "6-31g", 40),
    "h2o32_631gd": (lambda seed=0: water_grid(4, 4, 2, seed=seed), "6-31g(d)", 160),
    "h2o64_ccpvdz": (lambda seed=0: water_grid(4, 4, 4, seed=seed), "cc-pvdz", 320),
    "c20h42_631gd": (lambda seed=0: alkane(20, seed=seed), "6-31g(d)", 81),
    "h2o16_ccpvtz": (lambda seed=0: water_grid(4, 2, 2, seed=seed), "cc-pvtz", 80),
}
