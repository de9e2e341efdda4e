"""numpy restatement of OpenQP's MRSF-TDDFT sigma product (A - B) X in the
MO frame (the native path that routec_sig_iter replaces,
tdhf_mrsf_energy.F90:686-790): iatogen -> mrsfcbc -> int2_mrsf_data_t ->
mrsfmntoia + mrsfesum, for the singlet (kind 1) and triplet (kind 3) MRSF
response with every spin-pair-coupling scale equal to the HF scale.

Test infrastructure only (the checker of the device sigma); never on the
product path.  Matrices follow the reference's Fortran (column-major) index
order: M[mu, nu] here is M(mu, nu) there, 0-based.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

ISQ2 = 1.0 / np.sqrt(2.0)


def iatogen(pv, nbf, nocca, noccb):
    """iatogen (tdhf_lib.F90:480-498): the trial vector as the nbf x nbf
    matrix with block (1:nocca, noccb+1:nbf) filled column-major."""
    av = np.zeros((nbf, nbf))
    av[:nocca, noccb:] = np.asarray(pv).reshape(nbf - noccb, nocca).T
    return av


def mrsfcbc(va, vb, X, nocca, noccb, kind):
    """mrsfcbc (tdhf_mrsf_lib.F90:940-1273): the seven AO transition
    densities [bo2v, bo1v, bco1, bco2, o21v, co12, ball] of trial matrix X."""
    nbf = va.shape[0]
    lr1, lr2 = nocca - 2, nocca - 1  # 0-based HOMO-1, HOMO
    vir = slice(nocca, nbf)
    out = np.zeros((7, nbf, nbf))
    t = vb[:, vir] @ X[lr2, vir]
    out[0] += np.outer(va[:, lr2], t)                       # bo2v (:986-1005)
    t = vb[:, vir] @ X[lr1, vir]
    out[1] += np.outer(va[:, lr1], t)                       # bo1v (:1022-1034)
    if noccb > 0:
        t = va[:, :noccb] @ X[:noccb, lr1]
        out[2] += np.outer(t, vb[:, lr1])                   # bco1 (:1055-1065)
        t = va[:, :noccb] @ X[:noccb, lr2]
        out[3] += np.outer(t, vb[:, lr2])                   # bco2 (:1083-1093)
    t = vb[:, vir] @ X[lr2, vir]
    out[4] += np.outer(t, va[:, lr1])                       # o21v (:1100-1108)
    t = vb[:, vir] @ X[lr1, vir]
    out[4] -= np.outer(t, va[:, lr2])                       # (:1110-1118)
    if noccb > 0:
        t = va[:, :noccb] @ X[:noccb, lr1]
        out[5] += np.outer(vb[:, lr2], t)                   # co12 (:1125-1133)
        t = va[:, :noccb] @ X[:noccb, lr2]
        out[5] -= np.outer(vb[:, lr1], t)                   # (:1135-1143)
    ball = out[0] + out[1] + out[2] + out[3]                 # (:1146)
    if noccb > 0:
        t = vb[:, vir] @ X[:noccb, vir].T                   # (:1149-1157)
        ball += va[:, :noccb] @ t.T
    if kind == 1:                                            # (:1160-1167)
        ball += (np.outer(va[:, lr2], vb[:, lr1]) * X[lr2, lr1] + np.outer(va[:, lr1], vb[:, lr2]) * X[lr1, lr2]
                 + (np.outer(va[:, lr1], vb[:, lr1]) - np.outer(va[:, lr2], vb[:, lr2])) * X[lr1, lr1] * ISQ2)
    elif kind == 3:                                          # (:1168-1173)
        ball += (np.outer(va[:, lr1], vb[:, lr1]) + np.outer(va[:, lr2], vb[:, lr2])) * X[lr1, lr1] * ISQ2
    out[6] = ball
    return out


def mrsfmntoia(f, va, vb, nocca, noccb, kind):
    """mrsfmntoia (tdhf_mrsf_lib.F90:1463-1735): the seven digested AO
    matrices [ado2v, ado1v, adco1, adco2, ao21v, aco12, agdlr] back to the
    (1:nocca, noccb+1:nbf) block, packed column-major."""
    ado2v, ado1v, adco1, adco2, ao21v, aco12, agdlr = f
    nbf = va.shape[0]
    lr1, lr2 = nocca - 2, nocca - 1
    scr = va.T @ agdlr @ vb                                   # (:1515-1525)
    wrk = scr.copy()
    t = ado1v @ vb[:, lr2] + aco12 @ vb[:, lr1]               # (:1527-1535)
    if nocca > 2:
        wrk[:nocca - 2, lr2] += va[:, :nocca - 2].T @ t       # (:1541-1546)
    t = ado2v @ vb[:, lr1] - aco12 @ vb[:, lr2]               # (:1548-1556)
    if nocca > 2:
        wrk[:nocca - 2, lr1] += va[:, :nocca - 2].T @ t
    t = adco2.T @ va[:, lr1] + ao21v.T @ va[:, lr2]           # (:1566-1574)
    wrk[lr1, nocca:] += vb[:, nocca:].T @ t
    t = adco1.T @ va[:, lr2] - ao21v.T @ va[:, lr1]           # (:1582-1590)
    wrk[lr2, nocca:] += vb[:, nocca:].T @ t
    if kind == 1:                                             # (:1598-1605)
        wrk[lr1, lr1] = (scr[lr1, lr1] - scr[lr2, lr2]) * ISQ2
        wrk[lr2, lr2] = 0.0
    elif kind == 3:
        wrk[lr1, lr1] = (scr[lr1, lr1] + scr[lr2, lr2]) * ISQ2
        wrk[lr2, lr1] = wrk[lr1, lr2] = wrk[lr2, lr2] = 0.0
    return wrk[:nocca, noccb:].T.reshape(-1).copy()            # (:1607-1613)


def mrsfesum(X, fij, fab, nocca, noccb, kind):
    """mrsfesum (tdhf_mrsf_lib.F90:1918-2036): the orbital-energy part, packed
    like mrsfmntoia."""
    nbf = X.shape[0]
    lr1, lr2 = nocca - 2, nocca - 1
    scr = X.copy()
    scr[lr1, lr1] = 0.0
    scr[lr2, lr2] = 0.0
    tmp1 = np.zeros((nbf, nbf))
    tmp1[:nocca, noccb:] = scr[:nocca, noccb:] @ fab[noccb:, noccb:].T - fij[:nocca, :nocca] @ scr[:nocca, noccb:]
    xlr = X[lr1, lr1]
    p = -1.0 if kind == 1 else 1.0  # singlet / triplet signs of the HOMO terms (:1955-1995)
    w = np.zeros((nbf, nbf))
    w[:nocca, noccb:] = tmp1[:nocca, noccb:]
    w[lr1, noccb:] += fab[noccb:, lr1] * xlr * ISQ2
    w[lr2, noccb:] += p * fab[noccb:, lr2] * xlr * ISQ2
    w[:nocca, lr1] -= fij[:nocca, lr1] * xlr * ISQ2
    w[:nocca, lr2] -= p * fij[:nocca, lr2] * xlr * ISQ2
    dumn = (-fij[lr1, :nocca] @ scr[:nocca, lr1] - p * (fij[lr2, :nocca] @ scr[:nocca, lr2])
            + fab[lr1, noccb:] @ scr[lr1, noccb:] + p * (fab[lr2, noccb:] @ scr[lr2, noccb:]))
    w[lr1, lr1] = dumn * ISQ2 + xlr * (fab[lr1, lr1] + fab[lr2, lr2] - fij[lr1, lr1] - fij[lr2, lr2]) * 0.5
    if kind == 1:
        w[lr2, lr2] = 0.0
    else:
        w[lr2, lr1] = w[lr1, lr2] = w[lr2, lr2] = 0.0
    return w[:nocca, noccb:].T.reshape(-1).copy()


def sigma(bas, va, vb, fa, fb, nocca, noccb, kind, scale, bvec, cutoff=5e-11):
    """(A - B) X for the columns of bvec (ntrial x nv): the native MRSF sigma
    (tdhf_mrsf_energy.F90:686-790) with the integral pass of the oracle."""
    nbf = va.shape[0]
    bvec = np.asarray(bvec, np.float64).reshape(nocca * (nbf - noccb), -1, order="F")
    nv = bvec.shape[1]
    Xs = [iatogen(bvec[:, v], nbf, nocca, noccb) for v in range(nv)]
    d3 = np.zeros((nbf, nbf, 7, nv))  # C order [nu, mu, c, v] = Fortran d3(v, c, mu, nu)
    for v, X in enumerate(Xs):
        comps = mrsfcbc(va, vb, X, nocca, noccb, kind)
        for c in range(7):
            d3[:, :, c, v] = comps[c].T
    f3, _ = O.mrsf_int2(bas, d3, ncoul=4, se=scale, sc=scale, cutoff=cutoff)
    out = np.zeros((nocca * (nbf - noccb), nv))
    for v, X in enumerate(Xs):
        f = np.array([f3[:, :, c, v].T for c in range(7)])
        if kind == 3:                                   # (tdhf_mrsf_energy.F90:756-761)
            f[:6] = -f[:6]
        out[:, v] = mrsfmntoia(f, va, vb, nocca, noccb, kind) + mrsfesum(X, fa, fb, nocca, noccb, kind)
    return out
