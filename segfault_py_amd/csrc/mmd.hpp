// mmd.hpp -- McMurchie-Davidson building blocks (host+device):
// Hermite expansion coefficients E^{ij}_t and Coulomb auxiliary R_{tuv}.
//
// Reference counterparts: mmd.hermite_gauss_e (src/hf/core/core.f90:30-61)
// and mmd.coulomb_auxiliary_r (core.f90:64-97), there evaluated by
// exponential-cost recursion per call; here by the same recurrences run
// bottom-up once per primitive pair into small register tables.
//   E^{00}_0 = 1 (the exp(-mu X_AB^2) factor is carried by K of the pair),
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1},
//   E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1};
//   R^n_{000} = (-2p)^n F_n(p R_PC^2),
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X_PC R^{n+1}_{tuv} (same in u, v).
#pragma once
#include "eri_core.hpp"

namespace sf {

// E[i][j][t] for i <= IM, j <= JM, t <= i + j (entries with t > i+j are 0).
template <int IM, int JM>
SF_HD SF_INLINE void hermite_E(double ip2, double PA, double PB, double (&E)[IM + 1][JM + 1][IM + JM + 1])
{
    constexpr int NT = IM + JM + 1;
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int t = 0; t <= i + 1; ++t) {
            double v = PA * E[i][0][t];
            if (t > 0) v += ip2 * E[i][0][t - 1];
            if (t + 1 <= i) v += (t + 1) * E[i][0][t + 1];
            E[i + 1][0][t] = v;
        }
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j < JM; ++j)
#pragma unroll
            for (int t = 0; t <= i + j + 1; ++t) {
                double v = PB * E[i][j][t];
                if (t > 0) v += ip2 * E[i][j][t - 1];
                if (t + 1 <= i + j) v += (t + 1) * E[i][j][t + 1];
                E[i][j + 1][t] = v;
            }
}

// R[t][u][v] = R^0_{tuv} for t + u + v <= L, from F_0..F_L at T = p R_PC^2.
template <int L>
SF_HD SF_INLINE void coulomb_R(double p, double X, double Y, double Z, const double* F,
                               double (&R)[L + 1][L + 1][L + 1])
{
    // Rn[n][t][u][v]: build from n = L down to 0
    double Rn[L + 1][L + 1][L + 1][L + 1];
    double m2p = 1.0;
#pragma unroll
    for (int n = 0; n <= L; ++n) {
#pragma unroll
        for (int t = 0; t <= L; ++t)
#pragma unroll
            for (int u = 0; u <= L; ++u)
#pragma unroll
                for (int v = 0; v <= L; ++v) Rn[n][t][u][v] = 0.0;
        Rn[n][0][0][0] = m2p * F[n];
        m2p *= -2.0 * p;
    }
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
        // fill all t+u+v <= L-n from level n+1
#pragma unroll
        for (int s = 1; s <= L - n; ++s)
#pragma unroll
            for (int t = 0; t <= s; ++t)
#pragma unroll
                for (int u = 0; u <= s - t; ++u) {
                    const int v = s - t - u;
                    double r;
                    if (t > 0) {
                        r = X * Rn[n + 1][t - 1][u][v];
                        if (t > 1) r += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        r = Y * Rn[n + 1][t][u - 1][v];
                        if (u > 1) r += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        r = Z * Rn[n + 1][t][u][v - 1];
                        if (v > 1) r += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = r;
                }
    }
#pragma unroll
    for (int t = 0; t <= L; ++t)
#pragma unroll
        for (int u = 0; u <= L; ++u)
#pragma unroll
            for (int v = 0; v <= L; ++v) R[t][u][v] = Rn[0][t][u][v];
}

// S, T, V blocks (shell-normalised components, row-major NA x NB) of one
// shell pair over its primitive pairs and natom point charges Z at Cxyz
// (reference: mmd2.py:14-84; the kinetic formula of kinetic_3d, :31-42).
template <int LA, int LB>
SF_HD SF_INLINE void one_electron_shell_pair(const PrimPair* __restrict__ prim, int np,
                                             const double* A, const double* B, int natom,
                                             const double* __restrict__ Zc,
                                             const double* __restrict__ Cxyz, double* s, double* t,
                                             double* v)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB), L = LA + LB;
#pragma unroll
    for (int e = 0; e < NA * NB; ++e) s[e] = t[e] = v[e] = 0.0;
    for (int k = 0; k < np; ++k) {
        const PrimPair pp = prim[k];
        const double p = pp.p, eb = pp.eb;
        const double P[3] = {pp.Px, pp.Py, pp.Pz};
        double E[3][LA + 1][LB + 3][LA + LB + 3];
#pragma unroll
        for (int d = 0; d < 3; ++d) hermite_E<LA, LB + 2>(0.5 * pp.ip, P[d] - A[d], P[d] - B[d], E[d]);
        const double sp = kPi * pp.ip;
        const double pre_s = pp.K * sp * sqrt(sp);
#pragma unroll
        for (int fa = 0; fa < NA; ++fa)
#pragma unroll
            for (int fb = 0; fb < NB; ++fb) {
                double S1[3], T1[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int ad = cart_l(LA, fa, d), bd = cart_l(LB, fb, d);
                    S1[d] = E[d][ad][bd][0];
                    double tk = -2.0 * eb * eb * E[d][ad][bd + 2][0] + eb * (2 * bd + 1) * S1[d];
                    if (bd >= 2) tk -= 0.5 * bd * (bd - 1) * E[d][ad][bd - 2][0];
                    T1[d] = tk;
                }
                s[fa * NB + fb] += pre_s * S1[0] * S1[1] * S1[2];
                t[fa * NB + fb] += pre_s * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] +
                                            S1[0] * S1[1] * T1[2]);
            }
        for (int c = 0; c < natom; ++c) {
            const double PC[3] = {P[0] - Cxyz[3 * c], P[1] - Cxyz[3 * c + 1], P[2] - Cxyz[3 * c + 2]};
            double F[SF_BOYS_MTAB];
            boys_grid(L, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F);
            double R[L + 1][L + 1][L + 1];
            coulomb_R<L>(p, PC[0], PC[1], PC[2], F, R);
            const double pre_v = -Zc[c] * 2.0 * kPi * pp.ip * pp.K;
#pragma unroll
            for (int fa = 0; fa < NA; ++fa)
#pragma unroll
                for (int fb = 0; fb < NB; ++fb) {
                    const int ax = cart_l(LA, fa, 0), ay = cart_l(LA, fa, 1), az = cart_l(LA, fa, 2);
                    const int bx = cart_l(LB, fb, 0), by = cart_l(LB, fb, 1), bz = cart_l(LB, fb, 2);
                    double acc = 0.0;
#pragma unroll
                    for (int tt = 0; tt <= ax + bx; ++tt)
#pragma unroll
                        for (int uu = 0; uu <= ay + by; ++uu)
#pragma unroll
                            for (int vv = 0; vv <= az + bz; ++vv)
                                acc += E[0][ax][bx][tt] * E[1][ay][by][uu] * E[2][az][bz][vv] *
                                       R[tt][uu][vv];
                    v[fa * NB + fb] += pre_v * acc;
                }
        }
    }
}

// Nuclear gradient of the one-electron energy of one shell pair:
//   row[3c+k] = d/dR_{c,k} of  mult * sum_{ab} [P_ab (T_ab + V_ab) - W_ab S_ab]
// for every atom c (the pair's centres and every nucleus), with P, W the
// density and energy-weighted density blocks in the shell-normalised basis
// (row-major NA x NB) and mult = 2 for a != b (the transposed block), 1 else.
// d/dA of a primitive: 2a (a+1| - a_k (a-1|; S and T are translation
// invariant (d/dB = -d/dA); for the attraction to nucleus C,
// d/dC = -(d/dA + d/dB) (the Hellmann-Feynman term).
// Overlap and kinetic part (translation invariant): gA += d/dA, gB += d/dB.
template <int LA, int LB>
SF_HD SF_INLINE void one_electron_grad_st(const PrimPair* __restrict__ prim, int np,
                                          const double* A, const double* B,
                                          const double* __restrict__ Pb,
                                          const double* __restrict__ Wb, double mult, double* gA,
                                          double* gB)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    for (int k = 0; k < np; ++k) {
        const PrimPair pp = prim[k];
        const double eb = pp.eb, ea2 = 2.0 * (pp.p - eb);
        const double P[3] = {pp.Px, pp.Py, pp.Pz};
        double E[3][LA + 2][LB + 3][LA + LB + 4];
#pragma unroll
        for (int d = 0; d < 3; ++d) hermite_E<LA + 1, LB + 2>(0.5 * pp.ip, P[d] - A[d], P[d] - B[d], E[d]);
        const double sp = kPi * pp.ip;
        const double pre_s = pp.K * sp * sqrt(sp);
#pragma unroll
        for (int fa = 0; fa < NA; ++fa)
#pragma unroll
            for (int fb = 0; fb < NB; ++fb) {
                const double Pv = Pb[fa * NB + fb], Wv = Wb[fa * NB + fb];
                double S0[3], T0[3], Sd[3], Td[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int ad = cart_l(LA, fa, d), bd = cart_l(LB, fb, d);
                    // 1-D overlap and kinetic factors at (ad, bd) and their A-derivative
                    auto s1 = [&](int i) { return E[d][i][bd][0]; };
                    auto t1 = [&](int i) {
                        double t = -2.0 * eb * eb * E[d][i][bd + 2][0] + eb * (2 * bd + 1) * E[d][i][bd][0];
                        if (bd >= 2) t -= 0.5 * bd * (bd - 1) * E[d][i][bd >= 2 ? bd - 2 : 0][0];
                        return t;
                    };
                    S0[d] = s1(ad);
                    T0[d] = t1(ad);
                    Sd[d] = ea2 * s1(ad + 1) - (ad > 0 ? ad * s1(ad > 0 ? ad - 1 : 0) : 0.0);
                    Td[d] = ea2 * t1(ad + 1) - (ad > 0 ? ad * t1(ad > 0 ? ad - 1 : 0) : 0.0);
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
                    const double dS = pre_s * Sd[d] * S0[d1] * S0[d2];
                    const double dT = pre_s * (Td[d] * S0[d1] * S0[d2] + Sd[d] * T0[d1] * S0[d2] +
                                               Sd[d] * S0[d1] * T0[d2]);
                    const double v = mult * (Pv * dT - Wv * dS);
                    gA[d] += v;
                    gB[d] -= v;
                }
            }
    }
}

// Attraction to nucleus c (charge Zc[c] at Cxyz[3c..]): gA, gB, gC (set,
// not accumulated) = d/dA, d/dB, d/dC, with d/dC = -(d/dA + d/dB).
template <int LA, int LB>
SF_HD SF_INLINE void one_electron_grad_v(const PrimPair* __restrict__ prim, int np,
                                         const double* A, const double* B, int c,
                                         const double* __restrict__ Zc,
                                         const double* __restrict__ Cxyz,
                                         const double* __restrict__ Pb, double mult, double* gA,
                                         double* gB, double* gC)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB), L = LA + LB + 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) gA[d] = gB[d] = gC[d] = 0.0;
    for (int k = 0; k < np; ++k) {
        const PrimPair pp = prim[k];
        const double p = pp.p, eb = pp.eb, ea2 = 2.0 * (p - eb), eb2 = 2.0 * eb;
        const double P[3] = {pp.Px, pp.Py, pp.Pz};
        double E[3][LA + 2][LB + 2][LA + LB + 3];
#pragma unroll
        for (int d = 0; d < 3; ++d) hermite_E<LA + 1, LB + 1>(0.5 * pp.ip, P[d] - A[d], P[d] - B[d], E[d]);
        const double PC[3] = {P[0] - Cxyz[3 * c], P[1] - Cxyz[3 * c + 1], P[2] - Cxyz[3 * c + 2]};
        double F[SF_BOYS_MTAB];
        boys_grid(L, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F);
        double R[L + 1][L + 1][L + 1];
        coulomb_R<L>(p, PC[0], PC[1], PC[2], F, R);
        const double pre_v = -Zc[c] * 2.0 * kPi * pp.ip * pp.K;
#pragma unroll
        for (int fa = 0; fa < NA; ++fa)
#pragma unroll
            for (int fb = 0; fb < NB; ++fb) {
                const double Pv = mult * pre_v * Pb[fa * NB + fb];
                const int a3[3] = {cart_l(LA, fa, 0), cart_l(LA, fa, 1), cart_l(LA, fa, 2)};
                const int b3[3] = {cart_l(LB, fb, 0), cart_l(LB, fb, 1), cart_l(LB, fb, 2)};
                // V(a', b') for exponent triples a', b'
                auto vab = [&](int ax, int ay, int az, int bx, int by, int bz) {
                    double acc = 0.0;
                    for (int tt = 0; tt <= ax + bx; ++tt)
                        for (int uu = 0; uu <= ay + by; ++uu)
                            for (int vv = 0; vv <= az + bz; ++vv)
                                acc += E[0][ax][bx][tt] * E[1][ay][by][uu] * E[2][az][bz][vv] *
                                       R[tt][uu][vv];
                    return acc;
                };
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    int ap[3] = {a3[0], a3[1], a3[2]}, bp[3] = {b3[0], b3[1], b3[2]};
                    ap[d] += 1;
                    double dA = ea2 * vab(ap[0], ap[1], ap[2], b3[0], b3[1], b3[2]);
                    if (a3[d] > 0) {
                        ap[d] -= 2;
                        dA -= a3[d] * vab(ap[0], ap[1], ap[2], b3[0], b3[1], b3[2]);
                    }
                    bp[d] += 1;
                    double dB = eb2 * vab(a3[0], a3[1], a3[2], bp[0], bp[1], bp[2]);
                    if (b3[d] > 0) {
                        bp[d] -= 2;
                        dB -= b3[d] * vab(a3[0], a3[1], a3[2], bp[0], bp[1], bp[2]);
                    }
                    gA[d] += Pv * dA;
                    gB[d] += Pv * dB;
                    gC[d] -= Pv * (dA + dB);
                }
            }
    }
}

// Nuclear gradient of the one-electron energy of one shell pair:
//   row[3c+k] = d/dR_{c,k} of  mult * sum_{ab} [P_ab (T_ab + V_ab) - W_ab S_ab]
// for every atom c (the pair's centres and every nucleus), with P, W the
// density and energy-weighted density blocks in the shell-normalised basis
// (row-major NA x NB) and mult = 2 for a != b (the transposed block), 1 else.
// d/dA of a primitive: 2a (a+1| - a_k (a-1|; S and T are translation
// invariant (d/dB = -d/dA); for the attraction to nucleus C,
// d/dC = -(d/dA + d/dB) (the Hellmann-Feynman term).  The GPU computes the
// per-nucleus terms in parallel (k_grad_onee_v) and adds them in nucleus
// order to the overlap/kinetic part (k_grad_onee_st); this is the same
// computation on one thread, in the same order.
template <int LA, int LB>
SF_HD SF_INLINE void one_electron_grad_pair(const PrimPair* __restrict__ prim, int np,
                                            const double* A, const double* B, int atomA, int atomB,
                                            int natom, const double* __restrict__ Zc,
                                            const double* __restrict__ Cxyz,
                                            const double* __restrict__ Pb, const double* __restrict__ Wb,
                                            double mult, double* __restrict__ row)
{
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    one_electron_grad_st<LA, LB>(prim, np, A, B, Pb, Wb, mult, gA, gB);
    for (int c = 0; c < natom; ++c) {
        double vA[3], vB[3], vC[3];
        one_electron_grad_v<LA, LB>(prim, np, A, B, c, Zc, Cxyz, Pb, mult, vA, vB, vC);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            row[3 * c + d] = vC[d];
            gA[d] += vA[d];
            gB[d] += vB[d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        row[3 * atomA + d] += gA[d];
        row[3 * atomB + d] += gB[d];
    }
}

}  // namespace sf
