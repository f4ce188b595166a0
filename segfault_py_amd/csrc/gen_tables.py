"""Generate the numeric tables used by the HIP kernels (rys_tables.inc).

Run once in a container with mpmath; the generated file is committed.

1. Rys quadrature, n = 1..7 nodes (7 nodes are exact up to total angular
   momentum 13, i.e. every f-shell quartet): for X in [0, XMAX[n]) split into unit
   intervals, each root u_r(X) = t_r^2 and weight w_r(X) is a degree-11
   Chebyshev series on its interval (interpolation at the 12 Chebyshev
   points).  The exact values at the nodes come from Gautschi's Chebyshev
   algorithm on the Boys moments F_k(X) (k < 2n) followed by Golub-Welsch,
   all at 50 significant digits.  For X >= XMAX[n] the kernels use the
   large-X limit u_r = h_r^2 / X, w_r = W_r / sqrt(X), with h_r, W_r the
   positive nodes and weights of the 2n-point Gauss-Hermite rule (also
   computed here at 50 digits); XMAX[n] is where that limit is exact to
   < 2e-15 relative.
2. Boys function Taylor grid: F_m(k * 0.1) for k = 0..BOYS_NGRID-1,
   m = 0..BOYS_MTAB-1, for the expansion
   F_m(X) = sum_j F_{m+j}(X_k) (X_k - X)^j / j!.
"""
from __future__ import annotations

import os
import sys

import mpmath as mp

mp.mp.dps = 50

NMAX = 7
XMAX = {1: 40, 2: 48, 3: 54, 4: 60, 5: 66, 6: 72, 7: 78}
DEG = 11
BOYS_DX = 0.1
BOYS_NGRID = 401          # grid up to X = 40
BOYS_MTAB = 34            # m = 0..33: mmax 24 plus 9 Taylor terms


def boys_mp(m: int, X) -> mp.mpf:
    X = mp.mpf(X)
    if X == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, X) / (2 * X ** (m + mp.mpf(1) / 2))


def rys_mp(n: int, X):
    """Nodes u = t^2 (ascending) and weights for exp(-X t^2) on [0, 1]."""
    mu = [boys_mp(k, X) for k in range(2 * n)]
    alpha = [mu[1] / mu[0]]
    beta = [mu[0]]
    sig_m1 = [mp.mpf(0)] * (2 * n + 1)
    sig_0 = list(mu) + [mp.mpf(0)]
    for k in range(1, n):
        sig_k = [mp.mpf(0)] * (2 * n + 1)
        for l in range(k, 2 * n - k):
            sig_k[l] = sig_0[l + 1] - alpha[k - 1] * sig_0[l] - beta[k - 1] * sig_m1[l]
        alpha.append(sig_k[k + 1] / sig_k[k] - sig_0[k] / sig_0[k - 1])
        beta.append(sig_k[k] / sig_0[k - 1])
        sig_m1, sig_0 = sig_0, sig_k
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in idx], [beta[0] * Q[0, i] ** 2 for i in idx]


def hermite_half(n: int):
    """Positive nodes h_r (ascending) and weights of the 2n-point Gauss-Hermite rule."""
    m = 2 * n
    J = mp.matrix(m, m)
    for i in range(m - 1):
        J[i, i + 1] = J[i + 1, i] = mp.sqrt(mp.mpf(i + 1) / 2)
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0)
    return [p[0] for p in pairs], [p[1] for p in pairs]


def cheb_coefs(values_at_nodes):
    """Chebyshev coefficients from values at x_k = cos(pi (k+1/2)/(M+1))."""
    M = len(values_at_nodes) - 1
    out = []
    for j in range(M + 1):
        s = mp.mpf(0)
        for k in range(M + 1):
            s += values_at_nodes[k] * mp.cos(j * mp.pi * (k + mp.mpf(1) / 2) / (M + 1))
        c = 2 * s / (M + 1)
        if j == 0:
            c /= 2
        out.append(c)
    return out


def cheb_eval(c, s):
    """sum_j c_j T_j(s) (Clenshaw, as the kernels evaluate it)."""
    b1 = b2 = mp.mpf(0)
    for j in range(len(c) - 1, 0, -1):
        b1, b2 = c[j] + 2 * s * b1 - b2, b1
    return c[0] + s * b1 - b2


# the tables of n <= NCHECKED nodes were checked against mpmath by the tests
# when they were generated; larger n are checked here as well
NCHECKED = 5


def fmt(v) -> str:
    return mp.nstr(v, 20, min_fixed=-1, max_fixed=-1) if v != 0 else "0.0"


def main(path: str):
    lines = ["// Generated by gen_tables.py -- do not edit.", "// clang-format off", ""]
    lines.append(f"#define SF_RYS_DEG {DEG}")
    lines.append(f"#define SF_BOYS_DX {BOYS_DX!r}")
    lines.append(f"#define SF_BOYS_NGRID {BOYS_NGRID}")
    lines.append(f"#define SF_BOYS_MTAB {BOYS_MTAB}")
    lines.append(f"#define SF_RYS_NMAX {NMAX}")
    lines.append(f"SF_TABLE_QUAL int sf_rys_xmax[{NMAX + 1}] = {{0, " + ", ".join(str(XMAX[n]) for n in range(1, NMAX + 1)) + "};")
    M = DEG
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / (M + 1)) for k in range(M + 1)]
    for n in range(1, NMAX + 1):
        h, W = hermite_half(n)
        lines.append(f"SF_TABLE_QUAL double sf_herm_u{n}[{n}] = {{" + ", ".join(fmt(x * x) for x in h) + "};")
        lines.append(f"SF_TABLE_QUAL double sf_herm_w{n}[{n}] = {{" + ", ".join(fmt(x) for x in W) + "};")
        nint = XMAX[n]
        # layout: [interval][value: u_0..u_{n-1}, w_0..w_{n-1}][coef 0..DEG]
        lines.append(f"SF_TABLE_QUAL double sf_rys_cheb{n}[{nint * 2 * n * (DEG + 1)}] = {{")
        worst = mp.mpf(0)
        for it in range(nint):
            a, b = mp.mpf(it), mp.mpf(it + 1)
            vals = [rys_mp(n, a + (b - a) * (s + 1) / 2) for s in nodes]
            coefs = []
            for v in range(2 * n):
                series = [vals[k][0][v] if v < n else vals[k][1][v - n] for k in range(M + 1)]
                c = cheb_coefs(series)
                coefs.append(c)
                lines.append("  " + ", ".join(fmt(x) for x in c) + ",")
            if n > NCHECKED:
                # relative interpolation error between the interpolation points
                for s in (mp.mpf(-3) / 4, mp.mpf(-1) / 3, mp.mpf(0), mp.mpf(1) / 2, mp.mpf(7) / 8):
                    u, w = rys_mp(n, a + (b - a) * (s + 1) / 2)
                    for v in range(2 * n):
                        ref = u[v] if v < n else w[v - n]
                        worst = max(worst, abs(cheb_eval(coefs[v], s) - ref) / abs(ref))
            print(f"n={n} interval {it}", file=sys.stderr) if it % 20 == 0 else None
        lines.append("};")
        if n > NCHECKED:
            u, w = rys_mp(n, XMAX[n])
            lim = max(max(abs(h[r] ** 2 / XMAX[n] - u[r]) / u[r], abs(W[r] / mp.sqrt(XMAX[n]) - w[r]) / w[r])
                      for r in range(n))
            print(f"n={n}: interpolation error {mp.nstr(worst, 3)}, large-X limit error at X={XMAX[n]} "
                  f"{mp.nstr(lim, 3)}", file=sys.stderr)
            assert worst < 1e-14 and lim < 2e-15, (n, worst, lim)
    lines.append(f"SF_TABLE_QUAL double sf_boys_grid[{BOYS_NGRID * BOYS_MTAB}] = {{")
    for k in range(BOYS_NGRID):
        X = mp.mpf(k) / 10  # k * BOYS_DX exactly
        lines.append("  " + ", ".join(fmt(boys_mp(m, X)) for m in range(BOYS_MTAB)) + ",")
    lines.append("};")
    lines.append("// clang-format on")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "rys_tables.inc")
    main(out)
