// eri_gen.hpp -- contracted ERIs of one Cartesian component quartet with
// the angular momenta as run-time values (the f-shell quartet classes), and
// one element of a quartet's J/K digestion tiles from raw integral planes.
// Host+device code: the HIP kernels of k_eri_gen.hip use it on gfx950, and
// the CPU unit harness of the tests compiles it with g++.
//
// Reference: rys.eri_3d_rys / recur / shift (src/hf/core/core.f90:815-964)
// per component, as there; the node count is floor(L/2) + 1 <= 7 (exact),
// where the reference always takes 5 (exact only up to L = 9).
#pragma once
#include "eri_core.hpp"

namespace sf {

// Angular momenta and sizes of a quartet class for the run-time-L kernels:
// components per shell, integrals, Rys nodes, digestion tile offsets (X_ab,
// X_cd, Y_ac, Y_ad, Y_bc, Y_bd, as digest_quartet) and their total.
struct GenClass {
    int l[4];
    int n[4];
    int nint;
    int nr;
    int toff[6];
    int ntile;
};

SF_HD SF_INLINE GenClass gen_class(int la, int lb, int lc, int ld)
{
    GenClass g;
    const int l[4] = {la, lb, lc, ld};
    g.nint = 1;
    for (int s = 0; s < 4; ++s) {
        g.l[s] = l[s];
        g.n[s] = cart_n(l[s]);
        g.nint *= g.n[s];
    }
    g.nr = (la + lb + lc + ld) / 2 + 1;
    const int sz[6] = {g.n[0] * g.n[1], g.n[2] * g.n[3], g.n[0] * g.n[2],
                       g.n[0] * g.n[3], g.n[1] * g.n[2], g.n[1] * g.n[3]};
    g.toff[0] = 0;
    for (int t = 1; t < 6; ++t) g.toff[t] = g.toff[t - 1] + sz[t - 1];
    g.ntile = g.toff[5] + sz[5];
    return g;
}

// One primitive quartet of one Cartesian component quartet: exponents
// lc[12] = (lx, ly, lz) of a, b, c, d; the primitive pairs carry P, 1/p,
// K/p with the contraction coefficients folded in.
template <int N>
SF_HD SF_INLINE double gen_prim(const int* lc, const double* A, const double* C,
                                           const double* AB, const double* CD,
                                           const PrimPair& pb, const PrimPair& pk)
{
    const double p = pb.p, q = pk.p;
    const double rs = fast_rsqrt(p + q);
    const double ipq = rs * rs;
    const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
    const double Pv[3] = {pb.Px, pb.Py, pb.Pz};
    const double Qv[3] = {pk.Px, pk.Py, pk.Pz};
    const double X = p * q * ipq * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    double u[N], w[N];
    rys_roots<N>(X, u, w);
    const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
    double res = 0.0;
    for (int r = 0; r < N; ++r) {
        const double t2 = u[r];
        const double b00 = 0.5 * ipq * t2;
        const double b10 = 0.5 * pb.ip * (1.0 - q * ipq * t2);
        const double b01 = 0.5 * pk.ip * (1.0 - p * ipq * t2);
        double prod = pref * w[r];
        for (int k = 0; k < 3; ++k) {
            const int li = lc[k], lj = lc[3 + k], lk = lc[6 + k], ll = lc[9 + k];
            const int nn = li + lj, mm = lk + ll;
            const double c00 = (Pv[k] - A[k]) - q * ipq * PQ[k] * t2;
            const double c00p = (Qv[k] - C[k]) + p * ipq * PQ[k] * t2;
            // 2-D recurrence G(n, m), n <= nn, m <= mm (rys.recur, core.f90:815-874)
            double G[7][7];
            G[0][0] = 1.0;
            if (nn > 0) G[1][0] = c00;
            for (int n = 1; n < nn; ++n) G[n + 1][0] = n * b10 * G[n - 1][0] + c00 * G[n][0];
            for (int m = 0; m < mm; ++m)
                for (int n = 0; n <= nn; ++n) {
                    double g = c00p * G[n][m];
                    if (m > 0) g += m * b01 * G[n][m - 1];
                    if (n > 0) g += n * b00 * G[n - 1][m];
                    G[n][m + 1] = g;
                }
            // transfer to (li lj | lk ll) (rys.shift, core.f90:877-899)
            double s = 0.0, xkl = 1.0;
            for (int m = ll; m >= 0; --m) {
                double t = 0.0, xij = 1.0;
                for (int n = lj; n >= 0; --n) {
                    t += binom_c(lj, n) * xij * G[n + li][m + lk];
                    xij *= AB[k];
                }
                s += binom_c(ll, m) * xkl * t;
                xkl *= CD[k];
            }
            prod *= s;
        }
        res += prod;
    }
    return res;
}

// Contracted component quartet: primitive pairs of the bra (nb) and ket
// (nk), each ordered by its own Schwarz factor, loops stopping at the first
// primitive quartet with Qp_b Qp_k < tau (as eri_shell_quartet's screened
// loops; tau = 0 keeps every one).
template <int N>
SF_HD double gen_component(const int* lc, const double* A, const double* B, const double* C,
                                const double* D, const PrimPair* __restrict__ bra, int nb,
                                const PrimPair* __restrict__ ket, int nk,
                                const double* __restrict__ bq, const double* __restrict__ kq,
                                double tau)
{
    const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    const double CD[3] = {C[0] - D[0], C[1] - D[1], C[2] - D[2]};
    double acc = 0.0;
    for (int ik = 0; ik < nk; ++ik) {
        const double qk = kq ? kq[ik] : 1.0;
        if (kq && (nb == 0 || qk * bq[0] < tau)) break;
        const PrimPair pk = ket[ik];
        for (int ib = 0; ib < nb; ++ib) {
            if (kq && bq[ib] * qk < tau) break;
            acc += gen_prim<N>(lc, A, C, AB, CD, bra[ib], pk);
        }
    }
    return acc;
}

SF_HD SF_INLINE void gen_exps(const GenClass& g, int k, int* lc)
{
    int idx[4];
    idx[3] = k % g.n[3];
    k /= g.n[3];
    idx[2] = k % g.n[2];
    k /= g.n[2];
    idx[1] = k % g.n[1];
    idx[0] = k / g.n[1];
    for (int s = 0; s < 4; ++s)
        for (int e = 0; e < 3; ++e) lc[3 * s + e] = cart_l(g.l[s], idx[s], e);
}

// One Cartesian direction of (a_k b_k | c d) with the bra exponents a_k,
// b_k known only at run time: the 2-D recurrence over every row n <=
// LA + LB (compile-time bound), the ket transfer, and the bra transfer
// I = sum_n coef[n] G'(n, ...) with coef[n] = C(b_k, n - a_k) AB^(b_k - n +
// a_k) for a_k <= n <= a_k + b_k and 0 elsewhere -- so every array index is
// static and the tables stay in registers.
template <int LAB, int LC, int LD>
SF_HD SF_INLINE void dir_runtime(double g00, double c00, double c00p, double b10, double b01,
                                 double b00, const double (&coef)[LAB + 1], const double* CDp,
                                 double (&I)[LC + 1][LD + 1])
{
    constexpr int NN = LAB + 1, NM = LC + LD + 1;
    double G[NN][NM];
    G[0][0] = g00;
    if (NN > 1) G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < NN - 1; ++n) G[n + 1][0] = n * b10 * G[n - 1][0] + c00 * G[n][0];
    if (NM > 1) {
#pragma unroll
        for (int n = 0; n < NN; ++n) {
            G[n][1] = c00p * G[n][0];
            if (n > 0) G[n][1] += n * b00 * G[n - 1][0];
        }
#pragma unroll
        for (int m = 1; m < NM - 1; ++m)
#pragma unroll
            for (int n = 0; n < NN; ++n) {
                double g = m * b01 * G[n][m - 1] + c00p * G[n][m];
                if (n > 0) g += n * b00 * G[n - 1][m];
                G[n][m + 1] = g;
            }
    }
#pragma unroll
    for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) {
            double s = 0.0;
#pragma unroll
            for (int n = 0; n < NN; ++n) {
                double h = 0.0;
#pragma unroll
                for (int m = 0; m <= id; ++m) h += binom_c(id, m) * CDp[id - m] * G[n][ic + m];
                s += coef[n] * h;
            }
            I[ic][id] = s;
        }
}

// (fa fb | fc fd) for ONE bra component pair given at run time (fa of
// shell A, fb of shell B) and all ket components, accumulated into
// acc[NC*ND] -- the large f-shell classes, one wavefront per (64 quartets,
// bra component pair).  Primitive loops (ket outer, bra inner) and their
// screening as eri_bra_component; nodes from the global tables.
template <int LA, int LB, int LC, int LD>
SF_HD SF_INLINE void eri_bra_runtime(int fa, int fb, const PrimPair* __restrict__ bra, int nbra,
                                     const PrimPair* __restrict__ ket, int nket, const double* A,
                                     const double* B, const double* C, const double* D,
                                     double* __restrict__ acc, const double* __restrict__ bq,
                                     const double* __restrict__ kq, double tau_prim)
{
    constexpr int LAB = LA + LB, NC = cart_n(LC), ND = cart_n(LD);
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    double coef[3][LAB + 1], CDp[3][LD + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ak = cart_l(LA, fa, k), bk = cart_l(LB, fb, k);
        const double ab = A[k] - B[k], cd = C[k] - D[k];
#pragma unroll
        for (int n = 0; n <= LAB; ++n) {
            const int j = n - ak;
            double c = 0.0;
            if (j >= 0 && j <= bk) {
                c = binom_c(bk, j);
                for (int e = j; e < bk; ++e) c *= ab;
            }
            coef[k][n] = c;
        }
        CDp[k][0] = 1.0;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    const NodeTable nt = global_node_table<NR>();
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if (kq) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if (kq && bq[ib] * qk < tau_prim) break;
            const PrimPair pb = bra[ib];
            const double p = pb.p;
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double rs = fast_rsqrt(p + q);
            const double ipq = rs * rs;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = p * q * ipq * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                const double b00 = 0.5 * ipq * t2;
                const double b10 = 0.5 * pb.ip * (1.0 - q * ipq * t2);
                const double b01 = 0.5 * pk.ip * (1.0 - p * ipq * t2);
                double Ix[LC + 1][LD + 1], Iy[LC + 1][LD + 1], Iz[LC + 1][LD + 1];
                dir_runtime<LAB, LC, LD>(pref * w[r], PA[0] - q * ipq * PQ[0] * t2,
                                         QC[0] + p * ipq * PQ[0] * t2, b10, b01, b00, coef[0],
                                         CDp[0], Ix);
                dir_runtime<LAB, LC, LD>(1.0, PA[1] - q * ipq * PQ[1] * t2,
                                         QC[1] + p * ipq * PQ[1] * t2, b10, b01, b00, coef[1],
                                         CDp[1], Iy);
                dir_runtime<LAB, LC, LD>(1.0, PA[2] - q * ipq * PQ[2] * t2,
                                         QC[2] + p * ipq * PQ[2] * t2, b10, b01, b00, coef[2],
                                         CDp[2], Iz);
#pragma unroll
                for (int fc = 0; fc < NC; ++fc)
#pragma unroll
                    for (int fd = 0; fd < ND; ++fd)
                        acc[fc * ND + fd] += Ix[cart_l(LC, fc, 0)][cart_l(LD, fd, 0)] *
                                             Iy[cart_l(LC, fc, 1)][cart_l(LD, fd, 1)] *
                                             Iz[cart_l(LC, fc, 2)][cart_l(LD, fd, 2)];
            }
        }
    }
}

// Element e of digestion tile T (0..5: X_ab, X_cd, Y_ac, Y_ad, Y_bc, Y_bd;
// the layout and contraction of digest_quartet) of one quartet: integral k
// of the quartet is raw[k * vs]; fac is the degeneracy factor; P the
// density; qm the quartet's matrix positions.  Terms are added in the
// lexicographic order of (a, b, c, d), as digest_quartet does.
SF_HD SF_INLINE double gen_tile_element(const GenClass& g, int T, int e,
                                        const double* __restrict__ raw, long long vs, double fac,
                                        const double* __restrict__ P, const QuartetMap& qm)
{
    // output shells (o1, o2) of the tile; the density block P[c1][c2] (rows
    // of shell c1, columns of shell c2) multiplies the integrals
    const int o1 = (T == 0 || T == 2 || T == 3) ? 0 : (T == 1) ? 2 : 1;
    const int o2 = (T == 0) ? 1 : (T == 1) ? 3 : (T == 2 || T == 4) ? 2 : 3;
    const int c1 = (T == 0) ? 2 : (T == 1) ? 0 : (T == 2 || T == 3) ? 1 : 0;
    const int c2 = (T == 0 || T == 2 || T == 4) ? 3 : (T == 1) ? 1 : 2;
    const int x = e / g.n[o2], y = e % g.n[o2];
    const double scale = (T <= 1) ? 2.0 : 1.0;
    const double* Pr = P + qm.row[c1] + qm.col[c2];
    const int st[4] = {g.n[1] * g.n[2] * g.n[3], g.n[2] * g.n[3], g.n[3], 1};
    const int first = c1 < c2 ? c1 : c2, second = c1 < c2 ? c2 : c1;
    double t = 0.0;
    for (int u = 0; u < g.n[first]; ++u)
        for (int v = 0; v < g.n[second]; ++v) {
            int idx[4];
            idx[o1] = x;
            idx[o2] = y;
            idx[first] = u;
            idx[second] = v;
            const long long k = idx[0] * st[0] + idx[1] * st[1] + idx[2] * st[2] + idx[3];
            const double val = fac * raw[k * vs];
            t += scale * val * Pr[(long long)idx[c1] * qm.ld + idx[c2]];
        }
    return t;
}

}  // namespace sf
