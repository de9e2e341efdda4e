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
    vals = [v.strip().rstrip("f") for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
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


def _roots32(txt_tab, n, T):
    """FP32 nodes and weights exactly as the FP32 kernels evaluate them:
    degree-6 Horner in float32 on the same unit intervals, Hermite limit with
    float32 reciprocals beyond TASYM."""
    TAS, OFF32, TAB32, HO, HE = txt_tab
    T = np.float32(T)
    if T >= np.float32(TAS[n]):
        it, rs = np.float32(1) / T, np.float32(1) / np.sqrt(T)
        return (HE[HO[n]:HO[n] + n].astype(np.float32) * it, HE[HO[n] + n:HO[n] + 2 * n].astype(np.float32) * rs)
    k = int(T)
    s = np.float32(2) * (T - np.float32(k)) - np.float32(1)
    base = OFF32[n] + k * 2 * n * (DEG32 + 1)
    out = []
    for f in range(2 * n):
        v = np.float32(0)
        for c in TAB32[base + f * (DEG32 + 1): base + (f + 1) * (DEG32 + 1)][::-1]:
            v = np.float32(v * s + c)
        out.append(v)
    return np.array(out[:n], np.float32), np.array(out[n:], np.float32)


DEG32 = 6


def test_rys_fp32_moments(oracle):
    """The FP32 root table of the mixed-precision path: Boys moments to 2e-6
    relative in float32 arithmetic (every moment a sum of positive terms), for
    every root count and the whole T range, including both sides of TASYM."""
    txt = open(os.path.join(ROOT, "openqp_amd", "csrc", "rys_table.inc")).read()
    t32 = open(os.path.join(ROOT, "openqp_amd", "csrc", "rys_table_f32.inc")).read()
    tab = (_arr(txt, "RYS_TASYM").astype(int), _arr(t32, "RYS32_OFF").astype(int),
           _arr(t32, "RYS32_TABLE").astype(np.float32), _arr(txt, "RYS_HOFF").astype(int), _arr(txt, "RYS_HERM"))
    assert int(re.search(r"#define RYS32_DEG (\d+)", t32).group(1)) == DEG32
    rng = np.random.default_rng(8)
    worst = 0.0
    for n in range(1, 9):
        Ts = np.concatenate([rng.uniform(0, 100, 120), rng.uniform(0, 2, 30),
                             [0.0, 1e-6, tab[0][n] - 1e-3, tab[0][n] + 1e-3, float(n) + 0.5]])
        for T in Ts:
            x, w = _roots32(tab, n, T)
            F = oracle.boys(2 * n - 1, float(np.float32(T)))
            mom = np.array([np.sum(w.astype(np.float64) * x.astype(np.float64) ** k) for k in range(2 * n)])
            rel = np.max(np.abs(mom - F) / F)
            worst = max(worst, rel)
            assert rel < 2e-6, (n, T, rel)
            assert np.all(x > 0) and np.all(w > 0)
    print("worst FP32 moment error", worst)
