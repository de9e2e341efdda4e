"""Generate the Rys quadrature tables used by the device kernels.

For n roots and Boys argument T the Rys nodes x_i = t_i^2 and weights w_i
are the Gauss rule of the measure exp(-T x) x^(-1/2)/2 dx on [0, 1], whose
moments are the Boys functions F_k(T), k = 0..2n-1.  Here they are computed
in 60-digit arithmetic (mpmath): Boys moments by their convergent series,
Gautschi's Chebyshev algorithm for the three-term recurrence, and the
Golub-Welsch eigenproblem of the Jacobi matrix.

Each function x_i(T), w_i(T) is then fitted on unit intervals [k, k+1),
k < TASYM[n], by a degree-12 interpolant at Chebyshev points, stored as
monomial coefficients in s = 2 (T - k) - 1 (Horner on the device).  For
T >= TASYM[n] the kernels use the large-T limit x_i = r_i / T,
w_i = h_i / sqrt(T) with r_i, h_i the squared positive Gauss-Hermite nodes
(2n-point rule) and their weights; TASYM[n] is the smallest integer T at
which that limit agrees with the exact rule to 1e-16 relative, plus a
margin of 2 (found by bisection over integer T against ``rys_exact``).

Output: openqp_amd/csrc/rys_table.inc (hex-float literals, exact doubles).
Run: python openqp_amd/tools/gen_rys_table.py [--nmax 8] [--jobs 8]
"""
from __future__ import annotations

import argparse
import os
import sys
from multiprocessing import Pool

from mpmath import mp, mpf, exp, sqrt, pi, factorial, polyroots, hermite, cos

mp.dps = 60
DEG = 12
TASYM = {1: 41, 2: 49, 3: 56, 4: 62, 5: 68, 6: 74, 7: 80, 8: 85, 9: 91}


def boys_all(mmax, T):
    T = mpf(T)
    if T == 0:
        return [mpf(1) / (2 * m + 1) for m in range(mmax + 1)]
    m = mmax
    term = mpf(1) / (2 * m + 1)
    s = term
    k = 1
    while True:
        term *= 2 * T / (2 * m + 2 * k + 1)
        s += term
        k += 1
        if term < s * mpf(10) ** (-66):
            break
    eT = exp(-T)
    F = [None] * (mmax + 1)
    F[mmax] = eT * s
    for mm in range(mmax, 0, -1):
        F[mm - 1] = (2 * T * F[mm] + eT) / (2 * mm - 1)
    return F


def rys_exact(n, T):
    mu = boys_all(2 * n - 1, T)
    sig_prev = [mpf(0)] * (2 * n)
    sig = mu[:]
    alpha = [mu[1] / mu[0]]
    beta = [mu[0]]
    sigs = [sig_prev, sig]
    for k in range(1, n):
        sk = [mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sk[l] = sigs[-1][l + 1] - alpha[k - 1] * sigs[-1][l] - beta[k - 1] * sigs[-2][l]
        alpha.append(sk[k + 1] / sk[k] - sigs[-1][k] / sigs[-1][k - 1])
        beta.append(sk[k] / sigs[-1][k - 1])
        sigs.append(sk)
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in idx], [beta[0] * Q[0, i] ** 2 for i in idx]


def hermite_limit(n):
    """Squared positive nodes r_i and weights h_i of the 2n-point Gauss-Hermite rule."""
    N = 2 * n
    c_prev, c = [mpf(1)], [mpf(0), mpf(2)]
    for k in range(1, N):
        nxt = [mpf(0)] * (k + 2)
        for i, v in enumerate(c):
            nxt[i + 1] += 2 * v
        for i, v in enumerate(c_prev):
            nxt[i] -= 2 * k * v
        c_prev, c = c, nxt
    roots = polyroots(list(reversed(c)), maxsteps=400, extraprec=400)
    xs = sorted([mp.re(r) for r in roots if mp.re(r) > 0])
    ws = [2 ** (N - 1) * factorial(N) * sqrt(pi) / (N ** 2 * hermite(N - 1, x) ** 2) for x in xs]
    return [x * x for x in xs], ws


def cheb_to_monomial(c):
    """Chebyshev series coefficients -> monomial coefficients (mp)."""
    m = len(c) - 1
    T = [[mpf(1)], [mpf(0), mpf(1)]]
    for k in range(2, m + 1):
        a = [mpf(0)] + [2 * v for v in T[k - 1]]
        b = T[k - 2] + [mpf(0)] * (len(a) - len(T[k - 2]))
        T.append([a[i] - b[i] for i in range(len(a))])
    out = [mpf(0)] * (m + 1)
    for k in range(m + 1):
        for i, v in enumerate(T[k]):
            out[i] += c[k] * v
    return out


def fit_interval(args):
    n, k = args
    N = DEG + 1
    nodes = [cos(pi * (j + mpf(1) / 2) / N) for j in range(N)]
    vals = []
    for s in nodes:
        x, w = rys_exact(n, mpf(k) + (s + 1) / 2)
        vals.append(x + w)
    res = []
    for f in range(2 * n):
        c = []
        for kk in range(N):
            acc = sum(vals[j][f] * cos(pi * kk * (j + mpf(1) / 2) / N) for j in range(N))
            c.append(acc * (1 if kk == 0 else 2) / N)
        res.append([float(v) for v in cheb_to_monomial(c)])
    return n, k, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nmax", type=int, default=8)
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "csrc",
                                                  "rys_table.inc"))
    a = ap.parse_args()
    tasks = [(n, k) for n in range(1, a.nmax + 1) for k in range(TASYM[n])]
    with Pool(a.jobs) as pool:
        results = pool.map(fit_interval, tasks, chunksize=4)
    coef = {}
    for n, k, res in results:
        coef[(n, k)] = res
    lines = ["// Generated by openqp_amd/tools/gen_rys_table.py -- do not edit.",
             f"// Rys nodes/weights: degree-{DEG} monomial fits on unit T intervals, s = 2(T-k)-1.",
             f"#define RYS_NMAX {a.nmax}", f"#define RYS_DEG {DEG}"]
    offs, o = [], 0
    for n in range(1, a.nmax + 1):
        offs.append(o)
        o += TASYM[n] * 2 * n * (DEG + 1)
    lines.append("static constexpr int RYS_TASYM[RYS_NMAX + 1] = {0, " + ", ".join(str(TASYM[n]) for n in range(1, a.nmax + 1)) + "};")
    lines.append("static constexpr int RYS_OFF[RYS_NMAX + 1] = {0, " + ", ".join(str(v) for v in offs) + "};")
    lines.append(f"#define RYS_TABLE_SIZE {o}")
    lines.append("#ifdef OQPH_WANT_RYS_DATA")
    lines.append("static const double RYS_TABLE[RYS_TABLE_SIZE] = {")
    for n in range(1, a.nmax + 1):
        for k in range(TASYM[n]):
            for f in range(2 * n):
                lines.append("  " + ", ".join(float.hex(v) for v in coef[(n, k)][f]) + ",")
    lines.append("};")
    lines.append("#endif")
    # Hermite limit: r_i then h_i for each n
    hoffs, ho, hvals = [], 0, []
    for n in range(1, a.nmax + 1):
        r, h = hermite_limit(n)
        hoffs.append(ho)
        ho += 2 * n
        hvals.extend([float(v) for v in r] + [float(v) for v in h])
    lines.append("static constexpr int RYS_HOFF[RYS_NMAX + 1] = {0, " + ", ".join(str(v) for v in hoffs) + "};")
    lines.append("// small and indexed by compile-time constants in device code")
    lines.append(f"static constexpr double RYS_HERM[{ho}] = {{" + ", ".join(float.hex(v) for v in hvals) + "};")
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", a.out, "doubles", o)


if __name__ == "__main__":
    sys.exit(main())
