"""The device Rys table (openqp_amd/csrc/rys_table.inc) reproduces the Boys
moments sum_i w_i x_i^k = F_k(T), k < 2n, of the oracle's independent Boys
function: this pins every node and weight the kernels use.  CPU only."""
import os
import re

import numpy as np

from helpers import ROOT

DEG = 12


def _arr(txt, name):
    m = re.search(name + r"\[[^\]]*\] = \{(.*?)\};", txt, re.S)
    vals = [v.strip() for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    return np.array([float.fromhex(v) if "x" in v else float(v) for v in vals])


def _table():
    txt = open(os.path.join(ROOT, "openqp_amd", "csrc", "rys_table.inc")).read()
    return (_arr(txt, "RYS_TASYM").astype(int), _arr(txt, "RYS_OFF").astype(int), _arr(txt, "RYS_TABLE"),
            _arr(txt, "RYS_HOFF").astype(int), _arr(txt, "RYS_HERM"))


def _roots(tab, n, T):
    TAS, OFF, TAB, HO, HE = tab
    if T >= TAS[n]:
        return HE[HO[n]:HO[n] + n] / T, HE[HO[n] + n:HO[n] + 2 * n] / np.sqrt(T)
    k = int(T)
    s = 2 * (T - k) - 1
    base = OFF[n] + k * 2 * n * (DEG + 1)
    out = []
    for f in range(2 * n):
        v = 0.0
        for c in TAB[base + f * (DEG + 1): base + (f + 1) * (DEG + 1)][::-1]:
            v = v * s + c
        out.append(v)
    return np.array(out[:n]), np.array(out[n:])


def test_rys_moments(oracle):
    tab = _table()
    rng = np.random.default_rng(7)
    for n in range(1, 9):
        Ts = np.concatenate([rng.uniform(0, 100, 150), rng.uniform(0, 2, 40),
                             [0.0, 1e-12, tab[0][n] - 1e-9, tab[0][n] + 1e-9, float(n) + 0.5]])
        for T in Ts:
            x, w = _roots(tab, n, T)
            F = oracle.boys(2 * n - 1, T)
            mom = np.array([np.sum(w * x ** k) for k in range(2 * n)])
            assert np.max(np.abs(mom - F) / F) < 5e-15, (n, T)
            assert np.all(np.diff(x) > 0) and np.all(x > 0) and np.all(x < 1) and np.all(w > 0)
