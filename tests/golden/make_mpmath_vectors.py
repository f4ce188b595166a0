"""Golden vectors for the Boys function and Rys quadrature at 50 digits.

Boys: F_m(T) = gamma(m+1/2, T) / (2 T^(m+1/2)) by mpmath.gammainc.
Rys: n-node rule for exp(-X t^2) on [0, 1] in the variable u = t^2, from
the Boys moments by the Chebyshev algorithm and the Jacobi-matrix
eigenproblem; every rule is accepted only if it reproduces the 2n moments
F_0..F_{2n-1} to 1e-40 relative (which characterises the Gauss rule).

Writes boys_mpmath.json and rys_mpmath.json (values as decimal strings).
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))

T_POINTS = ["0", "1e-14", "1e-9", "1e-5", "0.003", "0.049", "0.05", "0.1", "0.37", "0.5", "0.99",
            "1", "2.5", "3.14159", "7.3", "10", "15.5", "19.99", "25", "29.9", "30", "33.3",
            "39.94", "39.96", "40", "41", "55", "66", "80", "150", "1000", "12345"]
X_POINTS = T_POINTS + ["0.8", "4.5", "47.9", "48.1", "53.9", "54.1", "59.9", "60.1", "65.9", "66.1",
                       "71.9", "72.1", "77.9", "78.1"]


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def rys(n, X):
    mu = [boys(k, X) for k in range(2 * n)]
    alpha, beta = [mu[1] / mu[0]], [mu[0]]
    s_m1 = [mp.mpf(0)] * (2 * n + 1)
    s_0 = list(mu) + [mp.mpf(0)]
    for k in range(1, n):
        s_k = [mp.mpf(0)] * (2 * n + 1)
        for l in range(k, 2 * n - k):
            s_k[l] = s_0[l + 1] - alpha[k - 1] * s_0[l] - beta[k - 1] * s_m1[l]
        alpha.append(s_k[k + 1] / s_k[k] - s_0[k] / s_0[k - 1])
        beta.append(s_k[k] / s_0[k - 1])
        s_m1, s_0 = s_0, s_k
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    u = [E[i] for i in idx]
    w = [beta[0] * Q[0, i] ** 2 for i in idx]
    for k in range(2 * n):
        got = mp.fsum(wi * ui ** k for ui, wi in zip(u, w))
        assert abs(got - mu[k]) <= mp.mpf("1e-40") * mu[k], (n, X, k)
    return u, w


def main():
    bo = {"mmax": 24, "points": []}
    for T in T_POINTS:
        bo["points"].append({"T": T, "F": [mp.nstr(boys(m, T), 30) for m in range(25)]})
    with open(os.path.join(HERE, "boys_mpmath.json"), "w") as fh:
        json.dump(bo, fh, indent=0)
    ry = {"points": []}
    for n in range(1, 8):
        for X in X_POINTS:
            u, w = rys(n, X)
            ry["points"].append({"n": n, "X": X, "u": [mp.nstr(v, 30) for v in u],
                                 "w": [mp.nstr(v, 30) for v in w]})
    with open(os.path.join(HERE, "rys_mpmath.json"), "w") as fh:
        json.dump(ry, fh, indent=0)


if __name__ == "__main__":
    main()
