// eri_core.hpp -- contracted shell-quartet ERIs by Rys quadrature, and the
// J/K digestion of one shell quartet.  Host+device code (see sf_math.hpp).
//
// Reference path restated (per Cartesian component there, per shell here):
//   rys.eri_rys     src/hf/core/core.f90:967-1002  (contraction)
//   rys.eri_3d_rys  core.f90:917-964              (rho, P, Q, X, 2 sqrt(rho/pi))
//   rys.recur       core.f90:815-874              (2-D recurrence)
//   rys.shift       core.f90:877-899              (horizontal transfer)
//   hf.py:300       G = J - K/2 from the N^4 tensor (digestion)
// Differences by design: the number of Rys nodes is floor(L/2)+1 (exact for
// the polynomial degree; the reference always uses 5), all Cartesian
// components of a shell quartet share one set of nodes and 2-D integrals,
// and the whole prefactor pi^3/(p q)^1.5 e^{...} 2 sqrt(rho/pi) is carried on
// the x direction.
#pragma once
#include <type_traits>

#include "sf_math.hpp"

namespace sf {

// Calls f(std::integral_constant<int, s>) for a runtime s in [I, N): turns a
// (wave-uniform) runtime index into a compile-time one.
template <int I, int N, typename F>
SF_HD SF_INLINE void static_switch(int s, F&& f)
{
    if constexpr (I < N) {
        if (s == I) f(std::integral_constant<int, I>{});
        else static_switch<I + 1, N>(s, f);
    }
}

// Calls f(std::integral_constant<int, i>) for i = I..N-1 in order.
template <int I, int N, typename F>
SF_HD SF_INLINE void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Cartesian component i of angular momentum L in the reference order
// (molecule.py:138-165): lx descending, then ly descending.
constexpr int cart_n(int L) { return (L + 1) * (L + 2) / 2; }
constexpr int cart_l(int L, int i, int axis)
{
    int c = 0;
    for (int lx = L; lx >= 0; --lx)
        for (int ly = L - lx; ly >= 0; --ly) {
            if (c == i) return axis == 0 ? lx : axis == 1 ? ly : L - lx - ly;
            ++c;
        }
    return 0;
}
constexpr double binom_c(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// One primitive pair of a shell pair (a, b): p = a_i + b_j, P = (a_i A + b_j B)/p,
// K = c_i c_j exp(-a_i b_j/p |A-B|^2) with the radial normalisation in c,
// and the per-pair constants the kernels would otherwise divide for.
struct PrimPair {
    double p, Px, Py, Pz, K;
    double ip;   // 1/p
    double Kp;   // K/p
    double eb;   // exponent b_j of the second shell
};

// Primitive pair i of an array given by any pointer type (generic, or the
// constant address space through which wave-uniform data is read with
// scalar loads); field by field, so no address-space-qualified copy.
template <typename Ptr>
SF_HD SF_INLINE PrimPair load_prim(Ptr ptr, int i)
{
    PrimPair r;
    r.p = ptr[i].p;
    r.Px = ptr[i].Px;
    r.Py = ptr[i].Py;
    r.Pz = ptr[i].Pz;
    r.K = ptr[i].K;
    r.ip = ptr[i].ip;
    r.Kp = ptr[i].Kp;
    r.eb = ptr[i].eb;
    return r;
}

// The primitive pairs of the ket pair lists transposed (one array per pair
// class, the kernels' per-lane ket reads): field f of primitive k of sorted
// pair j at t[(f * kmax + k) * n + j], fields p Px Py Pz ip Kp, then the
// Schwarz factor Qp and the charge S of each primitive pair.  Lanes holding
// consecutive kets read consecutive doubles (one cache line per 8 lanes,
// instead of one per lane from 64-byte records).  K and eb are not stored
// (the J/K kernels do not read them); the view returns them as NaN so that
// a use would show.
constexpr int kKetFields = 8;
constexpr int kKetQ = 6, kKetS = 7;
struct KetView {
    const double* t;  // array + j
    long long fs, n;  // field stride kmax * n, primitive stride n
    SF_HD SF_INLINE PrimPair operator[](int k) const
    {
        const double* e = t + (long long)k * n;
        PrimPair r;
        r.p = e[0];
        r.Px = e[fs];
        r.Py = e[2 * fs];
        r.Pz = e[3 * fs];
        r.ip = e[4 * fs];
        r.Kp = e[5 * fs];
        r.K = __builtin_nan("");
        r.eb = __builtin_nan("");
        return r;
    }
};
struct StridedView {
    const double* t;  // array + field offset + j
    long long n;
    SF_HD SF_INLINE double operator[](int k) const { return t[(long long)k * n]; }
};

// Where the four shells (a, b, c, d) of a quartet sit in an N x N matrix of
// their molecule: component x of shell k and component y of shell k' are at
// row[k] + x * ld + col[k'] + y.  For one molecule of nbf functions,
// row = foff * nbf, col = foff, ld = nbf.
struct QuartetMap {
    long long row[4];
    int col[4];
    int ld;
};

template <int LA, int LB, int LC, int LD>
struct QuartetShape {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int NR = L / 2 + 1;
    static constexpr int NA = cart_n(LA), NB = cart_n(LB), NC = cart_n(LC), ND = cart_n(LD);
    static constexpr int NINT = NA * NB * NC * ND;
    static constexpr int NAB = LA + LB + 1, NCD = LC + LD + 1;
    // digestion tiles: AB, CD, AC, AD, BC, BD
    static constexpr int T_AB = 0;
    static constexpr int T_CD = T_AB + NA * NB;
    static constexpr int T_AC = T_CD + NC * ND;
    static constexpr int T_AD = T_AC + NA * NC;
    static constexpr int T_BC = T_AD + NA * ND;
    static constexpr int T_BD = T_BC + NB * NC;
    static constexpr int NTILE = T_BD + NB * ND;
    // classes with more than 54 integrals per quartet run one wavefront per
    // component of A or per bra component pair (eri_bra_row /
    // eri_bra_component) and digest in a second kernel: measured on the
    // 20-water build, (pp|pp) (81 integrals, 256 VGPRs + 164 AGPRs as one
    // thread per quartet) is faster split, every class up to 54 integrals
    // is faster unsplit (the split recomputes nodes per component)
    static constexpr bool SPLIT = NINT > 54;
    // split classes whose NB*NC*ND integrals of one bra component of A fit
    // in registers run one wavefront per (64 quartets, component of A)
    // (eri_bra_row), the others one per (64 quartets, component pair)
    static constexpr bool SPLIT_ROWS = SPLIT && NB * NC * ND <= 54;
    static constexpr int NSLICE_ERI = SPLIT_ROWS ? NA : NA * NB;
    // nuclear gradient: classes with an s shell B (and at most 36 density
    // factors per quartet) run every component of A in one thread
    // (grad_all_a), the others one wavefront per bra component pair
    static constexpr bool GRAD_ALL_A = LB == 0 && NA * NC * ND <= 36;
    static constexpr int NSLICE_GRAD = GRAD_ALL_A ? 1 : NA * NB;
    static constexpr int tile_n1(int T) { return (T == 0 || T == 2 || T == 3) ? NA : (T == 1) ? NC : NB; }
    static constexpr int tile_n2(int T) { return (T == 0) ? NB : (T == 2 || T == 4) ? NC : ND; }
    static constexpr int tile_off(int T)
    {
        return T == 0 ? T_AB : T == 1 ? T_CD : T == 2 ? T_AC : T == 3 ? T_AD : T == 4 ? T_BC : T_BD;
    }
};

// One primitive quartet of a one-node class ((ss|ss): LA = 0, (ps|ss):
// LA = 1).  The Rys node u = F1/F0 only ever appears times the weight
// w = F0, so (ss|ss) = pref F0 and (p_k s|ss) = pref (PA_k F0 - q/(p+q)
// PQ_k F1), with no division.
template <int LA>
SF_HD SF_INLINE void one_node_term(const PrimPair& pb, const PrimPair& pk, const double* A,
                                   const NodeTable& nt, double* v)
{
    const double p = pb.p, q = pk.p;
    const double pq = p + q;
    const double rs = fast_rsqrt(pq);
    const double ipq = rs * rs;
    const double rho = p * q * ipq;
    const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
    const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
    double F0, F1;
    boys01_node(X, nt, F0, F1);
    if constexpr (LA == 0) {
        v[0] = pref * F0;
    } else {
        const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
        const double qi = q * ipq;
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = pref * (PA[k] * F0 - qi * PQ[k] * F1);
    }
}

// The one-node classes' primitive loops (ket outer, bra inner, as in
// eri_shell_quartet).  Two bra pairs per iteration, or prefetching the next
// pair, measured slower: the screened loops often stop early.
template <int LA, typename R, bool PS, typename BraPtr, typename BqPtr, typename KetPtr,
          typename KqPtr>
SF_HD SF_INLINE void eri_one_node(BraPtr bra, int nbra, KetPtr ket, int nket,
                                  const double* A, R* __restrict__ acc, const NodeTable& nt,
                                  BqPtr bq, KqPtr kq, double tau_prim)
{
    constexpr int NV = (LA == 0) ? 1 : 3;
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if constexpr (PS) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        for (int ib = 0; ib < nbra; ++ib) {
            if constexpr (PS) {
                if (bq[ib] * qk < tau_prim) break;
            }
            double v[NV];
            one_node_term<LA>(load_prim(bra, ib), pk, A, nt, v);
#pragma unroll
            for (int k = 0; k < NV; ++k) acc[k] += (R)v[k];
        }
    }
}

// The one-node classes on far quartets (every primitive quartet has
// X >= kFarX, where F_0 = sqrt(pi / (4X)) and F_1 = F_0 / (2X)): with
// S = pi^{3/2} K p^{-3/2} per primitive pair (bs / ks) and R = |P - Q|,
//   pref F_0 = S_b S_k / R,
//   pref (PA_k F_0 - q/(p+q) PQ_k F_1) = (S_b S_k / R) (PA_k - PQ_k / (2 p R^2)),
// the interaction of two point charges and of a charge with a dipole term:
// no Boys table, no 1/sqrt(p+q).  Same loops and primitive screening as
// eri_one_node.
template <int LA, typename R, bool PS, typename BraPtr, typename BsPtr, typename BqPtr,
          typename KetPtr, typename KsPtr, typename KqPtr>
SF_HD SF_INLINE void eri_one_node_far(BraPtr bra, BsPtr bs, int nbra, KetPtr ket, KsPtr ks,
                                      int nket, const double* A, R* __restrict__ acc, BqPtr bq,
                                      KqPtr kq, double tau_prim)
{
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if constexpr (PS) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const double Qx = ket[ik].Px, Qy = ket[ik].Py, Qz = ket[ik].Pz, sk = ks[ik];
        for (int ib = 0; ib < nbra; ++ib) {
            if constexpr (PS) {
                if (bq[ib] * qk < tau_prim) break;
            }
            const double PQ[3] = {bra[ib].Px - Qx, bra[ib].Py - Qy, bra[ib].Pz - Qz};
            const double ir = fast_rsqrt(PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double v0 = bs[ib] * sk * ir;
            if constexpr (LA == 0) {
                acc[0] += (R)v0;
            } else {
                const double c = 0.5 * bra[ib].ip * ir * ir;
                const double PA[3] = {bra[ib].Px - A[0], bra[ib].Py - A[1], bra[ib].Pz - A[2]};
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += (R)(v0 * (PA[k] - c * PQ[k]));
            }
        }
    }
}

// Accumulate (ab|cd) for all Cartesian components into acc[NINT] (not
// zeroed here).  Bra primitive pairs bra[0..nbra), ket pairs ket[0..nket).
// R is the type of the recurrences and of acc; X, nodes and weights are
// always evaluated in double, from the node table nt (global or a copy in
// LDS).  The ket primitive loop is the outer one: in the kernels each lane
// has its own ket pair (loaded once per quartet) while the bra pair is
// shared by the wavefront's lanes (one broadcast load per bra primitive).
// With PS (primitive screening), bq/kq are the primitive pairs' own Schwarz
// factors, sorted descending within each shell pair, and primitive quartets
// with bq * kq < tau_prim are skipped (both loops stop at the first one).
template <int LA, int LB, int LC, int LD, typename R, bool PS = false,
          typename BraPtr = const PrimPair*, typename BqPtr = const double*,
          typename KetPtr = const PrimPair*, typename KqPtr = const double*>
SF_HD SF_INLINE void eri_shell_quartet(BraPtr bra, int nbra, KetPtr ket, int nket,
                                       const double* A, const double* B, const double* C,
                                       const double* D, R* __restrict__ acc,
                                       const NodeTable& nt = global_node_table<QuartetShape<LA, LB, LC, LD>::NR>(),
                                       BqPtr bq = nullptr, KqPtr kq = nullptr,
                                       double tau_prim = 0.0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NR = S::NR, NAB = S::NAB, NCD = S::NCD;
    if constexpr (S::L <= 1) {
        static_assert(LB == 0 && LC == 0 && LD == 0, "the one-node classes are (ss|ss) and (ps|ss)");
        eri_one_node<LA, R, PS, BraPtr, BqPtr, KetPtr, KqPtr>(bra, nbra, ket, nket, A, acc, nt, bq,
                                                              kq, tau_prim);
        return;
    }
    R ABp[3][LB + 1], CDp[3][LD + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R ab = (R)(A[k] - B[k]), cd = (R)(C[k] - D[k]);
        ABp[k][0] = (R)1;
        CDp[k][0] = (R)1;
#pragma unroll
        for (int e = 1; e <= LB; ++e) ABp[k][e] = ABp[k][e - 1] * ab;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if constexpr (PS) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if constexpr (PS) {
                if (bq[ib] * qk < tau_prim) break;
            }
            const PrimPair pb = load_prim(bra, ib);
            const double p = pb.p;
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            // 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                // 2-D integrals G[n][m], n <= LA+LB, m <= LC+LD (rys.recur)
                R G[3][NAB][NCD];
                const R b00 = (R)(0.5 * ipq * t2);
                const R b10 = (R)(0.5 * pb.ip * (1.0 - q * ipq * t2));
                const R b01 = (R)(0.5 * pk.ip * (1.0 - p * ipq * t2));
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R c00 = (R)(PA[k] - q * ipq * PQ[k] * t2);
                    const R c00p = (R)(QC[k] + p * ipq * PQ[k] * t2);
                    G[k][0][0] = (k == 0) ? (R)(pref * w[r]) : (R)1;
                    if (NAB > 1) G[k][1][0] = c00 * G[k][0][0];
#pragma unroll
                    for (int n = 1; n < NAB - 1; ++n)
                        G[k][n + 1][0] = (R)n * b10 * G[k][n - 1][0] + c00 * G[k][n][0];
                    if (NCD > 1) {
#pragma unroll
                        for (int n = 0; n < NAB; ++n) {
                            G[k][n][1] = c00p * G[k][n][0];
                            if (n > 0) G[k][n][1] += (R)n * b00 * G[k][n - 1][0];
                        }
#pragma unroll
                        for (int m = 1; m < NCD - 1; ++m)
#pragma unroll
                            for (int n = 0; n < NAB; ++n) {
                                R g = (R)m * b01 * G[k][n][m - 1] + c00p * G[k][n][m];
                                if (n > 0) g += (R)n * b00 * G[k][n - 1][m];
                                G[k][n][m + 1] = g;
                            }
                    }
                }
                // horizontal transfer (rys.shift): ket then bra
                R H[3][NAB][LC + 1][LD + 1];
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int n = 0; n < NAB; ++n)
#pragma unroll
                        for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
                            for (int id = 0; id <= LD; ++id) {
                                R s = (R)0;
#pragma unroll
                                for (int m = 0; m <= id; ++m)
                                    s += (R)binom_c(id, m) * CDp[k][id - m] * G[k][n][ic + m];
                                H[k][n][ic][id] = s;
                            }
                R I[3][LA + 1][LB + 1][LC + 1][LD + 1];
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
                        for (int ib2 = 0; ib2 <= LB; ++ib2)
#pragma unroll
                            for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
                                for (int id = 0; id <= LD; ++id) {
                                    R s = (R)0;
#pragma unroll
                                    for (int n = 0; n <= ib2; ++n)
                                        s += (R)binom_c(ib2, n) * ABp[k][ib2 - n] * H[k][ia + n][ic][id];
                                    I[k][ia][ib2][ic][id] = s;
                                }
#pragma unroll
                for (int fa = 0; fa < S::NA; ++fa)
#pragma unroll
                    for (int fb = 0; fb < S::NB; ++fb)
#pragma unroll
                        for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                            for (int fd = 0; fd < S::ND; ++fd) {
                                const int idx = ((fa * S::NB + fb) * S::NC + fc) * S::ND + fd;
                                acc[idx] += I[0][cart_l(LA, fa, 0)][cart_l(LB, fb, 0)][cart_l(LC, fc, 0)][cart_l(LD, fd, 0)] *
                                            I[1][cart_l(LA, fa, 1)][cart_l(LB, fb, 1)][cart_l(LC, fc, 1)][cart_l(LD, fd, 1)] *
                                            I[2][cart_l(LA, fa, 2)][cart_l(LB, fb, 2)][cart_l(LC, fc, 2)][cart_l(LD, fd, 2)];
                            }
            }
        }
    }
}

// One Cartesian direction of (a_k b_k | c d) for fixed bra exponents
// AK, BK and all ket exponents ic <= LC, id <= LD: 2-D recurrence up to
// row AK+BK, ket transfer on rows AK..AK+BK, bra transfer to I[ic][id].
template <int AK, int BK, int LC, int LD, typename R>
SF_HD SF_INLINE void dir_component(R g00, R c00, R c00p, R b10, R b01, R b00, R ab,
                                   const R* CDp, R (&I)[LC + 1][LD + 1])
{
    constexpr int NN = AK + BK + 1, NM = LC + LD + 1;
    R G[NN][NM];
    G[0][0] = g00;
    if (NN > 1) G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < NN - 1; ++n) G[n + 1][0] = (R)n * b10 * G[n - 1][0] + c00 * G[n][0];
    if (NM > 1) {
#pragma unroll
        for (int n = 0; n < NN; ++n) {
            G[n][1] = c00p * G[n][0];
            if (n > 0) G[n][1] += (R)n * b00 * G[n - 1][0];
        }
#pragma unroll
        for (int m = 1; m < NM - 1; ++m)
#pragma unroll
            for (int n = 0; n < NN; ++n) {
                R g = (R)m * b01 * G[n][m - 1] + c00p * G[n][m];
                if (n > 0) g += (R)n * b00 * G[n - 1][m];
                G[n][m + 1] = g;
            }
    }
    R H[BK + 1][LC + 1][LD + 1];
#pragma unroll
    for (int nn = 0; nn <= BK; ++nn)
#pragma unroll
        for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
            for (int id = 0; id <= LD; ++id) {
                R s = (R)0;
#pragma unroll
                for (int m = 0; m <= id; ++m)
                    s += (R)binom_c(id, m) * CDp[id - m] * G[AK + nn][ic + m];
                H[nn][ic][id] = s;
            }
    R abp[BK + 1];
    abp[0] = (R)1;
#pragma unroll
    for (int e = 1; e <= BK; ++e) abp[e] = abp[e - 1] * ab;
#pragma unroll
    for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) {
            R s = (R)0;
#pragma unroll
            for (int nn = 0; nn <= BK; ++nn) s += (R)binom_c(BK, nn) * abp[BK - nn] * H[nn][ic][id];
            I[ic][id] = s;
        }
}

// (fa fb | fc fd) for ONE bra component pair (FA of shell A, FB of shell B)
// and all ket components, accumulated into acc[NC*ND].  Used where the full
// class does not fit in registers: one wavefront per bra component pair.
template <int LA, int LB, int LC, int LD, int FA, int FB, typename R, bool PS = false,
          typename KetPtr = const PrimPair*, typename KqPtr = const double*>
SF_HD SF_INLINE void eri_bra_component(const PrimPair* __restrict__ bra, int nbra,
                                       KetPtr ket, int nket,
                                       const double* A, const double* B, const double* C,
                                       const double* D, R* __restrict__ acc,
                                       const NodeTable& nt = global_node_table<QuartetShape<LA, LB, LC, LD>::NR>(),
                                       const double* __restrict__ bq = nullptr,
                                       KqPtr kq = nullptr,
                                       double tau_prim = 0.0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NR = S::NR;
    constexpr int ax = cart_l(LA, FA, 0), ay = cart_l(LA, FA, 1), az = cart_l(LA, FA, 2);
    constexpr int bx = cart_l(LB, FB, 0), by = cart_l(LB, FB, 1), bz = cart_l(LB, FB, 2);
    R CDp[3][LD + 1];
    R AB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R cd = (R)(C[k] - D[k]);
        AB[k] = (R)(A[k] - B[k]);
        CDp[k][0] = (R)1;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if constexpr (PS) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if constexpr (PS) {
                if (bq[ib] * qk < tau_prim) break;
            }
            const PrimPair pb = bra[ib];
            const double p = pb.p;
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                const R b00 = (R)(0.5 * ipq * t2);
                const R b10 = (R)(0.5 * pb.ip * (1.0 - q * ipq * t2));
                const R b01 = (R)(0.5 * pk.ip * (1.0 - p * ipq * t2));
                R c00[3], c00p[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c00[k] = (R)(PA[k] - q * ipq * PQ[k] * t2);
                    c00p[k] = (R)(QC[k] + p * ipq * PQ[k] * t2);
                }
                R Ix[LC + 1][LD + 1], Iy[LC + 1][LD + 1], Iz[LC + 1][LD + 1];
                dir_component<ax, bx, LC, LD, R>((R)(pref * w[r]), c00[0], c00p[0], b10, b01, b00,
                                                 AB[0], CDp[0], Ix);
                dir_component<ay, by, LC, LD, R>((R)1, c00[1], c00p[1], b10, b01, b00, AB[1],
                                                 CDp[1], Iy);
                dir_component<az, bz, LC, LD, R>((R)1, c00[2], c00p[2], b10, b01, b00, AB[2],
                                                 CDp[2], Iz);
#pragma unroll
                for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                    for (int fd = 0; fd < S::ND; ++fd)
                        acc[fc * S::ND + fd] += Ix[cart_l(LC, fc, 0)][cart_l(LD, fd, 0)] *
                                                Iy[cart_l(LC, fc, 1)][cart_l(LD, fd, 1)] *
                                                Iz[cart_l(LC, fc, 2)][cart_l(LD, fd, 2)];
            }
        }
    }
}

// One Cartesian direction for fixed bra exponent AK and every BK <= LB:
// I[bk][ic][id] = I(AK, bk, ic, id).  The 2-D recurrence rows up to AK+bk
// and the transfers are the same operations, in the same order, as
// dir_component<AK, bk> (identical values up to the compiler's choice of
// fused multiply-adds); here the recurrence and the ket transfer are done
// once for all bk.
template <int AK, int LB, int LC, int LD, typename R>
SF_HD SF_INLINE void dir_row(R g00, R c00, R c00p, R b10, R b01, R b00, R ab, const R* CDp,
                             R (&I)[LB + 1][LC + 1][LD + 1])
{
    constexpr int NN = AK + LB + 1, NM = LC + LD + 1;
    R G[NN][NM];
    G[0][0] = g00;
    if (NN > 1) G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < NN - 1; ++n) G[n + 1][0] = (R)n * b10 * G[n - 1][0] + c00 * G[n][0];
    if (NM > 1) {
#pragma unroll
        for (int n = 0; n < NN; ++n) {
            G[n][1] = c00p * G[n][0];
            if (n > 0) G[n][1] += (R)n * b00 * G[n - 1][0];
        }
#pragma unroll
        for (int m = 1; m < NM - 1; ++m)
#pragma unroll
            for (int n = 0; n < NN; ++n) {
                R g = (R)m * b01 * G[n][m - 1] + c00p * G[n][m];
                if (n > 0) g += (R)n * b00 * G[n - 1][m];
                G[n][m + 1] = g;
            }
    }
    R H[LB + 1][LC + 1][LD + 1];
#pragma unroll
    for (int nn = 0; nn <= LB; ++nn)
#pragma unroll
        for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
            for (int id = 0; id <= LD; ++id) {
                R s = (R)0;
#pragma unroll
                for (int m = 0; m <= id; ++m)
                    s += (R)binom_c(id, m) * CDp[id - m] * G[AK + nn][ic + m];
                H[nn][ic][id] = s;
            }
    R abp[LB + 1];
    abp[0] = (R)1;
#pragma unroll
    for (int e = 1; e <= LB; ++e) abp[e] = abp[e - 1] * ab;
#pragma unroll
    for (int bk = 0; bk <= LB; ++bk)
#pragma unroll
        for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
            for (int id = 0; id <= LD; ++id) {
                R s = (R)0;
#pragma unroll
                for (int nn = 0; nn <= bk; ++nn) s += (R)binom_c(bk, nn) * abp[bk - nn] * H[nn][ic][id];
                I[bk][ic][id] = s;
            }
}

// (fa fb | fc fd) for ONE bra component FA of shell A, every component fb
// of shell B and every ket component, accumulated into acc[NB*NC*ND]
// (index (fb*NC + fc)*ND + fd).  The same sums as eri_bra_component<FA, fb>
// for each fb, with the nodes, the 2-D recurrence and the ket transfer
// evaluated once per FA instead of once per (FA, fb).
template <int LA, int LB, int LC, int LD, int FA, typename R, bool PS = false,
          typename KetPtr = const PrimPair*, typename KqPtr = const double*>
SF_HD SF_INLINE void eri_bra_row(const PrimPair* __restrict__ bra, int nbra,
                                 KetPtr ket, int nket,
                                 const double* A, const double* B, const double* C,
                                 const double* D, R* __restrict__ acc,
                                 const NodeTable& nt = global_node_table<QuartetShape<LA, LB, LC, LD>::NR>(),
                                 const double* __restrict__ bq = nullptr,
                                 KqPtr kq = nullptr,
                                 double tau_prim = 0.0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NR = S::NR;
    constexpr int ax = cart_l(LA, FA, 0), ay = cart_l(LA, FA, 1), az = cart_l(LA, FA, 2);
    R CDp[3][LD + 1];
    R AB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R cd = (R)(C[k] - D[k]);
        AB[k] = (R)(A[k] - B[k]);
        CDp[k][0] = (R)1;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if constexpr (PS) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if constexpr (PS) {
                if (bq[ib] * qk < tau_prim) break;
            }
            const PrimPair pb = bra[ib];
            const double p = pb.p;
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                const R b00 = (R)(0.5 * ipq * t2);
                const R b10 = (R)(0.5 * pb.ip * (1.0 - q * ipq * t2));
                const R b01 = (R)(0.5 * pk.ip * (1.0 - p * ipq * t2));
                R c00[3], c00p[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c00[k] = (R)(PA[k] - q * ipq * PQ[k] * t2);
                    c00p[k] = (R)(QC[k] + p * ipq * PQ[k] * t2);
                }
                R Ix[LB + 1][LC + 1][LD + 1], Iy[LB + 1][LC + 1][LD + 1], Iz[LB + 1][LC + 1][LD + 1];
                dir_row<ax, LB, LC, LD, R>((R)(pref * w[r]), c00[0], c00p[0], b10, b01, b00, AB[0],
                                           CDp[0], Ix);
                dir_row<ay, LB, LC, LD, R>((R)1, c00[1], c00p[1], b10, b01, b00, AB[1], CDp[1], Iy);
                dir_row<az, LB, LC, LD, R>((R)1, c00[2], c00p[2], b10, b01, b00, AB[2], CDp[2], Iz);
#pragma unroll
                for (int fb = 0; fb < S::NB; ++fb)
#pragma unroll
                    for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                        for (int fd = 0; fd < S::ND; ++fd)
                            acc[(fb * S::NC + fc) * S::ND + fd] +=
                                Ix[cart_l(LB, fb, 0)][cart_l(LC, fc, 0)][cart_l(LD, fd, 0)] *
                                Iy[cart_l(LB, fb, 1)][cart_l(LC, fc, 1)][cart_l(LD, fd, 1)] *
                                Iz[cart_l(LB, fb, 2)][cart_l(LC, fc, 2)][cart_l(LD, fd, 2)];
            }
        }
    }
}

// ---------------------------------------------------------- gradients ----
// One Cartesian direction of the 2-D integrals around bra exponents (AK, BK)
// for the nuclear derivatives d/dA, d/dB, d/dC:
//   I0[ic][id]  = I(AK,   BK,   ic, id), ic <= LC+1,
//   IAp / IAm   = I(AK+-1, BK,   ic, id), ic <= LC  (IAm unused when AK = 0),
//   IBp / IBm   = I(AK,   BK+-1, ic, id), ic <= LC  (IBm unused when BK = 0).
template <int AK, int BK, int LC, int LD, typename R>
struct DirDeriv {
    R I0[LC + 2][LD + 1];
    R IAp[LC + 1][LD + 1], IAm[LC + 1][LD + 1];
    R IBp[LC + 1][LD + 1], IBm[LC + 1][LD + 1];
};

template <int AK, int BK, int LC, int LD, typename R>
SF_HD SF_INLINE void dir_deriv(R g00, R c00, R c00p, R b10, R b01, R b00, R ab, const R* CDp,
                               DirDeriv<AK, BK, LC, LD, R>& o)
{
    constexpr int NN = AK + BK + 2, NM = LC + LD + 2;
    R G[NN][NM];
    G[0][0] = g00;
    G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < NN - 1; ++n) G[n + 1][0] = (R)n * b10 * G[n - 1][0] + c00 * G[n][0];
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        G[n][1] = c00p * G[n][0];
        if (n > 0) G[n][1] += (R)n * b00 * G[n - 1][0];
    }
#pragma unroll
    for (int m = 1; m < NM - 1; ++m)
#pragma unroll
        for (int n = 0; n < NN; ++n) {
            R g = (R)m * b01 * G[n][m - 1] + c00p * G[n][m];
            if (n > 0) g += (R)n * b00 * G[n - 1][m];
            G[n][m + 1] = g;
        }
    // ket transfer on every row n, ic <= LC+1
    R H[NN][LC + 2][LD + 1];
#pragma unroll
    for (int n = 0; n < NN; ++n)
#pragma unroll
        for (int ic = 0; ic <= LC + 1; ++ic)
#pragma unroll
            for (int id = 0; id <= LD; ++id) {
                R s = (R)0;
#pragma unroll
                for (int m = 0; m <= id; ++m)
                    if (ic + m < NM) s += (R)binom_c(id, m) * CDp[id - m] * G[n][ic + m];
                H[n][ic][id] = s;
            }
    R abp[BK + 2];
    abp[0] = (R)1;
#pragma unroll
    for (int e = 1; e <= BK + 1; ++e) abp[e] = abp[e - 1] * ab;
    // bra transfer T(ia, ib) = sum_n C(ib,n) AB^(ib-n) H[ia+n]
#pragma unroll
    for (int ic = 0; ic <= LC + 1; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) {
            R s = (R)0;
#pragma unroll
            for (int n = 0; n <= BK; ++n) s += (R)binom_c(BK, n) * abp[BK - n] * H[AK + n][ic][id];
            o.I0[ic][id] = s;
        }
#pragma unroll
    for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) {
            R sp = (R)0, sm = (R)0, tp = (R)0, tm = (R)0;
#pragma unroll
            for (int n = 0; n <= BK; ++n) {
                sp += (R)binom_c(BK, n) * abp[BK - n] * H[AK + 1 + n][ic][id];
                if (AK > 0) sm += (R)binom_c(BK, n) * abp[BK - n] * H[(AK > 0 ? AK - 1 : 0) + n][ic][id];
            }
#pragma unroll
            for (int n = 0; n <= BK + 1; ++n) tp += (R)binom_c(BK + 1, n) * abp[BK + 1 - n] * H[AK + n][ic][id];
            if (BK > 0) {
#pragma unroll
                for (int n = 0; n <= (BK > 0 ? BK - 1 : 0); ++n)
                    tm += (R)binom_c(BK - 1, n) * abp[BK - 1 - n] * H[AK + n][ic][id];
            }
            o.IAp[ic][id] = sp;
            o.IAm[ic][id] = sm;
            o.IBp[ic][id] = tp;
            o.IBm[ic][id] = tm;
        }
}

// Nuclear-derivative contraction for ONE bra component pair (FA, FB) of a
// shell quartet: g[0..8] += sum_{fc,fd} D[fc*ND+fd] * d(fa fb|fc fd)/d(A, B, C)
// (x, y, z each), with d/dA of a primitive 2a (a+1 b| - a_k (a-1 b|, and the
// same for B (exponent b) and C (exponent c); the D-derivative follows from
// translational invariance, -(dA + dB + dC).  D is the density factor of
// the component quartet (see k_grad_split).
template <int LA, int LB, int LC, int LD, int FA, int FB, typename R>
SF_HD SF_INLINE void grad_bra_component(const PrimPair* __restrict__ bra, int nbra,
                                        const PrimPair* __restrict__ ket, int nket,
                                        const double* A, const double* B, const double* C,
                                        const double* D, const double* __restrict__ Dfac,
                                        double* __restrict__ g,
                                        const NodeTable& nt = global_node_table<(QuartetShape<LA, LB, LC, LD>::L + 1) / 2 + 1>(),
                                        const double* __restrict__ bq = nullptr,
                                        const double* __restrict__ kq = nullptr,
                                        double tau_prim = 0.0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NR = (S::L + 1) / 2 + 1;
    constexpr int ax = cart_l(LA, FA, 0), ay = cart_l(LA, FA, 1), az = cart_l(LA, FA, 2);
    constexpr int bx = cart_l(LB, FB, 0), by = cart_l(LB, FB, 1), bz = cart_l(LB, FB, 2);
    R CDp[3][LD + 1];
    R AB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R cd = (R)(C[k] - D[k]);
        AB[k] = (R)(A[k] - B[k]);
        CDp[k][0] = (R)1;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    // primitive screening as in eri_shell_quartet when bq/kq are given:
    // both loops stop at the first primitive quartet with bq * kq < tau_prim
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if (kq) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const R tc = (R)(2.0 * (q - pk.eb));
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if (kq && bq[ib] * qk < tau_prim) break;
            const PrimPair pb = bra[ib];
            const double p = pb.p;
            const R ta = (R)(2.0 * (p - pb.eb)), tb = (R)(2.0 * pb.eb);
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                const R b00 = (R)(0.5 * ipq * t2);
                const R b10 = (R)(0.5 * pb.ip * (1.0 - q * ipq * t2));
                const R b01 = (R)(0.5 * pk.ip * (1.0 - p * ipq * t2));
                R c00[3], c00p[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c00[k] = (R)(PA[k] - q * ipq * PQ[k] * t2);
                    c00p[k] = (R)(QC[k] + p * ipq * PQ[k] * t2);
                }
                DirDeriv<ax, bx, LC, LD, R> Dx;
                DirDeriv<ay, by, LC, LD, R> Dy;
                DirDeriv<az, bz, LC, LD, R> Dz;
                dir_deriv<ax, bx, LC, LD, R>((R)(pref * w[r]), c00[0], c00p[0], b10, b01, b00, AB[0],
                                             CDp[0], Dx);
                dir_deriv<ay, by, LC, LD, R>((R)1, c00[1], c00p[1], b10, b01, b00, AB[1], CDp[1], Dy);
                dir_deriv<az, bz, LC, LD, R>((R)1, c00[2], c00p[2], b10, b01, b00, AB[2], CDp[2], Dz);
#pragma unroll
                for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                    for (int fd = 0; fd < S::ND; ++fd) {
                        const int cx = cart_l(LC, fc, 0), cy = cart_l(LC, fc, 1), cz = cart_l(LC, fc, 2);
                        const int dx = cart_l(LD, fd, 0), dy = cart_l(LD, fd, 1), dz = cart_l(LD, fd, 2);
                        const R f = (R)Dfac[fc * S::ND + fd];
                        const R x0 = Dx.I0[cx][dx], y0 = Dy.I0[cy][dy], z0 = Dz.I0[cz][dz];
                        // d/dA
                        const R dax = ta * Dx.IAp[cx][dx] - (R)ax * Dx.IAm[cx][dx];
                        const R day = ta * Dy.IAp[cy][dy] - (R)ay * Dy.IAm[cy][dy];
                        const R daz = ta * Dz.IAp[cz][dz] - (R)az * Dz.IAm[cz][dz];
                        // d/dB
                        const R dbx = tb * Dx.IBp[cx][dx] - (R)bx * Dx.IBm[cx][dx];
                        const R dby = tb * Dy.IBp[cy][dy] - (R)by * Dy.IBm[cy][dy];
                        const R dbz = tb * Dz.IBp[cz][dz] - (R)bz * Dz.IBm[cz][dz];
                        // d/dC (ket component exponent c_k)
                        const R dcx = tc * Dx.I0[cx + 1][dx] - (cx > 0 ? (R)cx * Dx.I0[cx > 0 ? cx - 1 : 0][dx] : (R)0);
                        const R dcy = tc * Dy.I0[cy + 1][dy] - (cy > 0 ? (R)cy * Dy.I0[cy > 0 ? cy - 1 : 0][dy] : (R)0);
                        const R dcz = tc * Dz.I0[cz + 1][dz] - (cz > 0 ? (R)cz * Dz.I0[cz > 0 ? cz - 1 : 0][dz] : (R)0);
                        g[0] += (double)(f * dax * y0 * z0);
                        g[1] += (double)(f * x0 * day * z0);
                        g[2] += (double)(f * x0 * y0 * daz);
                        g[3] += (double)(f * dbx * y0 * z0);
                        g[4] += (double)(f * x0 * dby * z0);
                        g[5] += (double)(f * x0 * y0 * dbz);
                        g[6] += (double)(f * dcx * y0 * z0);
                        g[7] += (double)(f * x0 * dcy * z0);
                        g[8] += (double)(f * x0 * y0 * dcz);
                    }
            }
        }
    }
}

// The same derivatives as grad_bra_component for every component FA of A
// at once, for an s shell B (LB = 0): the nodes and weights of a primitive
// quartet are evaluated once instead of once per FA.  Dfac[FA*NC*ND + ...]
// holds each component's density factors; g[0..8] += all of them.
template <int LA, int LB, int LC, int LD, typename R>
SF_HD SF_INLINE void grad_all_a(const PrimPair* __restrict__ bra, int nbra,
                                const PrimPair* __restrict__ ket, int nket, const double* A,
                                const double* B, const double* C, const double* D,
                                const double* __restrict__ Dfac, double* __restrict__ g,
                                const NodeTable& nt, const double* __restrict__ bq = nullptr,
                                const double* __restrict__ kq = nullptr, double tau_prim = 0.0)
{
    static_assert(LB == 0, "grad_all_a: shell B must be an s shell");
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NR = (S::L + 1) / 2 + 1;
    constexpr int NK = S::NC * S::ND;
    R CDp[3][LD + 1];
    R AB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R cd = (R)(C[k] - D[k]);
        AB[k] = (R)(A[k] - B[k]);
        CDp[k][0] = (R)1;
#pragma unroll
        for (int e = 1; e <= LD; ++e) CDp[k][e] = CDp[k][e - 1] * cd;
    }
    // primitive screening as in eri_shell_quartet when bq/kq are given:
    // both loops stop at the first primitive quartet with bq * kq < tau_prim
    for (int ik = 0; ik < nket; ++ik) {
        double qk = 0.0;
        if (kq) {
            qk = kq[ik];
            if (nbra == 0 || qk * bq[0] < tau_prim) break;
        }
        const PrimPair pk = ket[ik];
        const double q = pk.p;
        const R tc = (R)(2.0 * (q - pk.eb));
        const double QC[3] = {pk.Px - C[0], pk.Py - C[1], pk.Pz - C[2]};
        for (int ib = 0; ib < nbra; ++ib) {
            if (kq && bq[ib] * qk < tau_prim) break;
            const PrimPair pb = bra[ib];
            const double p = pb.p;
            const R ta = (R)(2.0 * (p - pb.eb)), tb = (R)(2.0 * pb.eb);
            const double PA[3] = {pb.Px - A[0], pb.Py - A[1], pb.Pz - A[2]};
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double u[NR], w[NR];
            rys_nodes_tab<NR>(X, nt, u, w);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double t2 = u[r];
                const R b00 = (R)(0.5 * ipq * t2);
                const R b10 = (R)(0.5 * pb.ip * (1.0 - q * ipq * t2));
                const R b01 = (R)(0.5 * pk.ip * (1.0 - p * ipq * t2));
                R c00[3], c00p[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c00[k] = (R)(PA[k] - q * ipq * PQ[k] * t2);
                    c00p[k] = (R)(QC[k] + p * ipq * PQ[k] * t2);
                }
                const R g00x = (R)(pref * w[r]);
                static_for<0, S::NA>([&](auto fac) {
                    constexpr int FA = decltype(fac)::value;
                    constexpr int ax = cart_l(LA, FA, 0), ay = cart_l(LA, FA, 1), az = cart_l(LA, FA, 2);
                    DirDeriv<ax, 0, LC, LD, R> Dx;
                    DirDeriv<ay, 0, LC, LD, R> Dy;
                    DirDeriv<az, 0, LC, LD, R> Dz;
                    dir_deriv<ax, 0, LC, LD, R>(g00x, c00[0], c00p[0], b10, b01, b00, AB[0], CDp[0], Dx);
                    dir_deriv<ay, 0, LC, LD, R>((R)1, c00[1], c00p[1], b10, b01, b00, AB[1], CDp[1], Dy);
                    dir_deriv<az, 0, LC, LD, R>((R)1, c00[2], c00p[2], b10, b01, b00, AB[2], CDp[2], Dz);
#pragma unroll
                    for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                        for (int fd = 0; fd < S::ND; ++fd) {
                            const int cx = cart_l(LC, fc, 0), cy = cart_l(LC, fc, 1), cz = cart_l(LC, fc, 2);
                            const int dx = cart_l(LD, fd, 0), dy = cart_l(LD, fd, 1), dz = cart_l(LD, fd, 2);
                            const R f = (R)Dfac[FA * NK + fc * S::ND + fd];
                            const R x0 = Dx.I0[cx][dx], y0 = Dy.I0[cy][dy], z0 = Dz.I0[cz][dz];
                            const R dax = ta * Dx.IAp[cx][dx] - (R)ax * Dx.IAm[cx][dx];
                            const R day = ta * Dy.IAp[cy][dy] - (R)ay * Dy.IAm[cy][dy];
                            const R daz = ta * Dz.IAp[cz][dz] - (R)az * Dz.IAm[cz][dz];
                            const R dbx = tb * Dx.IBp[cx][dx];
                            const R dby = tb * Dy.IBp[cy][dy];
                            const R dbz = tb * Dz.IBp[cz][dz];
                            const R dcx = tc * Dx.I0[cx + 1][dx] - (cx > 0 ? (R)cx * Dx.I0[cx > 0 ? cx - 1 : 0][dx] : (R)0);
                            const R dcy = tc * Dy.I0[cy + 1][dy] - (cy > 0 ? (R)cy * Dy.I0[cy > 0 ? cy - 1 : 0][dy] : (R)0);
                            const R dcz = tc * Dz.I0[cz + 1][dz] - (cz > 0 ? (R)cz * Dz.I0[cz > 0 ? cz - 1 : 0][dz] : (R)0);
                            g[0] += (double)(f * dax * y0 * z0);
                            g[1] += (double)(f * x0 * day * z0);
                            g[2] += (double)(f * x0 * y0 * daz);
                            g[3] += (double)(f * dbx * y0 * z0);
                            g[4] += (double)(f * x0 * dby * z0);
                            g[5] += (double)(f * x0 * y0 * dbz);
                            g[6] += (double)(f * dcx * y0 * z0);
                            g[7] += (double)(f * x0 * dcy * z0);
                            g[8] += (double)(f * x0 * y0 * dcz);
                        }
                });
            }
        }
    }
}

// One digestion tile of a shell quartet from its raw integrals v (plane
// stride vs): T = 0 X_ab, 1 X_cd, 2 Y_ac, 3 Y_ad, 4 Y_bc, 5 Y_bd (see
// digest_quartet).  fac is the degeneracy factor.
// Geometry of digestion tile T (0..5: AB, CD, AC, AD, BC, BD) of a class:
// N1 x N2 outputs, contracted against the M1 x M2 density block whose rows
// and columns belong to shells RS and CS (0 a, 1 b, 2 c, 3 d); RI is the
// integral index (0 a, 1 b, 2 c) that selects the tile row.
//   T0 X_ab += 2 v P_cd, T1 X_cd += 2 v P_ab, T2 Y_ac += v P_bd,
//   T3 Y_ad += v P_bc, T4 Y_bc += v P_ad, T5 Y_bd += v P_ac.
template <int LA, int LB, int LC, int LD, int T>
struct TileGeom {
    using S = QuartetShape<LA, LB, LC, LD>;
    static constexpr int N1 = (T == 0 || T == 2 || T == 3) ? S::NA : (T == 1) ? S::NC : S::NB;
    static constexpr int N2 = (T == 0) ? S::NB : (T == 2 || T == 4) ? S::NC : S::ND;
    static constexpr int M1 = (T == 0) ? S::NC : (T == 1) ? S::NA : (T == 2 || T == 3) ? S::NB : S::NA;
    static constexpr int M2 = (T == 0) ? S::ND : (T == 1) ? S::NB : (T == 2) ? S::ND : (T == 3) ? S::NC : (T == 4) ? S::ND : S::NC;
    static constexpr int RS = (T == 0) ? 2 : (T == 1) ? 0 : (T == 2 || T == 3) ? 1 : 0;
    static constexpr int CS = (T == 0) ? 3 : (T == 1) ? 1 : (T == 2) ? 3 : (T == 3) ? 2 : (T == 4) ? 3 : 2;
    static constexpr int RI = (T == 0 || T == 2 || T == 3) ? 0 : (T == 1) ? 2 : 1;
};

template <int LA, int LB, int LC, int LD, int T>
SF_HD SF_INLINE void load_tile_density(const double* __restrict__ P, const QuartetMap& qm,
                                       double (&Pm)[TileGeom<LA, LB, LC, LD, T>::M1]
                                                    [TileGeom<LA, LB, LC, LD, T>::M2])
{
    using G = TileGeom<LA, LB, LC, LD, T>;
    const double* Pr = P + qm.row[G::RS] + qm.col[G::CS];
#pragma unroll
    for (int x = 0; x < G::M1; ++x)
#pragma unroll
        for (int y = 0; y < G::M2; ++y) Pm[x][y] = Pr[(long long)x * qm.ld + y];
}

// Row r of tile T: row[N2] = the outputs whose RI index is r.  Each output
// sums its terms in the lexicographic order of (a, b, c, d), so a tile built
// row by row equals the tile built at once bit for bit.  r may be a runtime
// (wavefront-uniform) value; it only offsets addresses.
template <int LA, int LB, int LC, int LD, int T>
SF_HD SF_INLINE void digest_tile_row(const double* __restrict__ v, long long vs, double fac,
                                     const double (&Pm)[TileGeom<LA, LB, LC, LD, T>::M1]
                                                       [TileGeom<LA, LB, LC, LD, T>::M2],
                                     int r, double* __restrict__ row)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    using G = TileGeom<LA, LB, LC, LD, T>;
#pragma unroll
    for (int e = 0; e < G::N2; ++e) row[e] = 0.0;
    if constexpr (G::RI == 0) {           // T0, T2, T3: a = r
        const double* va = v + (long long)r * (S::NB * S::NC * S::ND) * vs;
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int c = 0; c < S::NC; ++c)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const double x = fac * va[(long long)((b * S::NC + c) * S::ND + d) * vs];
                    if constexpr (T == 0) row[b] += 2.0 * x * Pm[c][d];
                    else if constexpr (T == 2) row[c] += x * Pm[b][d];
                    else row[d] += x * Pm[b][c];
                }
    } else if constexpr (G::RI == 2) {    // T1: c = r
#pragma unroll
        for (int a = 0; a < S::NA; ++a)
#pragma unroll
            for (int b = 0; b < S::NB; ++b)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const long long k = (long long)(((a * S::NB + b) * S::NC) * S::ND + d) + (long long)r * S::ND;
                    const double x = fac * v[k * vs];
                    row[d] += 2.0 * x * Pm[a][b];
                }
    } else {                               // T4, T5: b = r
#pragma unroll
        for (int a = 0; a < S::NA; ++a)
#pragma unroll
            for (int c = 0; c < S::NC; ++c)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const long long k = (long long)((a * S::NB * S::NC + c) * S::ND + d) +
                                        (long long)r * (S::NC * S::ND);
                    const double x = fac * v[k * vs];
                    if constexpr (T == 4) row[c] += x * Pm[a][d];
                    else row[d] += x * Pm[a][c];
                }
    }
}

// Whole tile T in one pass over the integrals (each output sums its terms
// in the same (a, b, c, d) order as digest_tile_row, so the two agree bit
// for bit); faster than row by row where the class has many quartets.
template <int LA, int LB, int LC, int LD, int T>
SF_HD SF_INLINE void digest_tile(const double* __restrict__ v, long long vs, double fac,
                                 const double* __restrict__ P, const QuartetMap& qm,
                                 double* __restrict__ t)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    using G = TileGeom<LA, LB, LC, LD, T>;
    constexpr int N2 = G::N2;
    double Pm[G::M1][G::M2];
    load_tile_density<LA, LB, LC, LD, T>(P, qm, Pm);
#pragma unroll
    for (int e = 0; e < G::N1 * N2; ++e) t[e] = 0.0;
#pragma unroll
    for (int a = 0; a < S::NA; ++a)
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int c = 0; c < S::NC; ++c)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const double x = fac * v[(long long)(((a * S::NB + b) * S::NC + c) * S::ND + d) * vs];
                    if constexpr (T == 0) t[a * N2 + b] += 2.0 * x * Pm[c][d];
                    else if constexpr (T == 1) t[c * N2 + d] += 2.0 * x * Pm[a][b];
                    else if constexpr (T == 2) t[a * N2 + c] += x * Pm[b][d];
                    else if constexpr (T == 3) t[a * N2 + d] += x * Pm[b][c];
                    else if constexpr (T == 4) t[b * N2 + c] += x * Pm[a][d];
                    else t[b * N2 + d] += x * Pm[a][c];
                }
}

// Doubles one lane holds to digest tile T: its N1 x N2 outputs and the
// M1 x M2 density block they contract with.
template <int LA, int LB, int LC, int LD>
constexpr int tile_lane_doubles(int T)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    const int out = S::tile_n1(T) * S::tile_n2(T);
    const int dens = T == 0 ? S::NC * S::ND : T == 1 ? S::NA * S::NB : T == 2 ? S::NB * S::ND
                   : T == 3 ? S::NB * S::NC : T == 4 ? S::NA * S::ND : S::NA * S::NC;
    return out + dens;
}

// The six tiles of a split class cut into contiguous groups of at most
// kDigestBudget doubles per lane (a tile above the budget is a group of its
// own): one wavefront digests one group in one pass over the integrals, so
// the raw planes are read once per group instead of once per tile.
// Measured on the 20-water build (FP64, ms per build): one tile per pass
// 6.40-6.52, budget 48: 6.32-6.35, 72: 6.35-6.37, 120: 6.07-6.10, 160:
// 6.08-6.16, 200: 6.19-6.25, 256: 6.14-6.33 (registers spill above 120).
constexpr int kDigestBudget = 120;
template <int LA, int LB, int LC, int LD>
struct TileGroups {
    struct Cuts {
        int n;
        int b[7];
    };
    static constexpr Cuts cuts()
    {
        Cuts c{0, {0, 0, 0, 0, 0, 0, 0}};
        int used = 0;
        for (int T = 0; T < 6; ++T) {
            const int d = tile_lane_doubles<LA, LB, LC, LD>(T);
            if (T > 0 && used + d > kDigestBudget) {
                c.b[++c.n] = T;
                used = 0;
            }
            used += d;
        }
        c.b[++c.n] = 6;
        return c;
    }
    static constexpr int count = cuts().n;
    static constexpr int begin(int g) { return cuts().b[g]; }
    static constexpr int end(int g) { return cuts().b[g + 1]; }
};

// Tiles TB..TE-1 of one quartet in one pass over its integrals: each tile
// adds its terms in the same order and with the same expressions as
// digest_tile, so the results equal digest_tile's bit for bit.
template <int LA, int LB, int LC, int LD, int TB, int TE>
SF_HD SF_INLINE void digest_tile_group(const double* __restrict__ v, long long vs, double fac,
                                       const double* __restrict__ P, const QuartetMap& qm,
                                       double* const (&dst)[6])
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr auto aoff = [](int T) {
        int o = 0;
        for (int u = TB; u < T; ++u) o += S::tile_n1(u) * S::tile_n2(u);
        return o;
    };
    constexpr auto poff = [](int T) {
        int o = 0;
        for (int u = TB; u < T; ++u)
            o += tile_lane_doubles<LA, LB, LC, LD>(u) - S::tile_n1(u) * S::tile_n2(u);
        return o;
    };
    constexpr int NACC = aoff(TE), NPM = poff(TE);
    double acc[NACC];
    double pm[NPM];
#pragma unroll
    for (int e = 0; e < NACC; ++e) acc[e] = 0.0;
    static_for<TB, TE>([&](auto tc) {
        constexpr int T = decltype(tc)::value;
        using G = TileGeom<LA, LB, LC, LD, T>;
        double Pm[G::M1][G::M2];
        load_tile_density<LA, LB, LC, LD, T>(P, qm, Pm);
#pragma unroll
        for (int x = 0; x < G::M1; ++x)
#pragma unroll
            for (int y = 0; y < G::M2; ++y) pm[poff(T) + x * G::M2 + y] = Pm[x][y];
    });
#pragma unroll
    for (int a = 0; a < S::NA; ++a)
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int c = 0; c < S::NC; ++c)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const double x = fac * v[(long long)(((a * S::NB + b) * S::NC + c) * S::ND + d) * vs];
                    static_for<TB, TE>([&](auto tc) {
                        constexpr int T = decltype(tc)::value;
                        using G = TileGeom<LA, LB, LC, LD, T>;
                        constexpr int N2 = G::N2, M2 = G::M2, A0 = aoff(T), P0 = poff(T);
                        if constexpr (T == 0) acc[A0 + a * N2 + b] += 2.0 * x * pm[P0 + c * M2 + d];
                        else if constexpr (T == 1) acc[A0 + c * N2 + d] += 2.0 * x * pm[P0 + a * M2 + b];
                        else if constexpr (T == 2) acc[A0 + a * N2 + c] += x * pm[P0 + b * M2 + d];
                        else if constexpr (T == 3) acc[A0 + a * N2 + d] += x * pm[P0 + b * M2 + c];
                        else if constexpr (T == 4) acc[A0 + b * N2 + c] += x * pm[P0 + a * M2 + d];
                        else acc[A0 + b * N2 + d] += x * pm[P0 + a * M2 + c];
                    });
                }
    static_for<TB, TE>([&](auto tc) {
        constexpr int T = decltype(tc)::value;
        constexpr int NE = S::tile_n1(T) * S::tile_n2(T), A0 = aoff(T);
#pragma unroll
        for (int e = 0; e < NE; ++e) dst[T][e] = acc[A0 + e];
    });
}

// J/K digestion of one shell quartet.  v[NINT] are the integrals in the
// shell-normalised basis times the degeneracy factor; P is the density in
// the same basis (row-major, ld = nbf).  Writes tiles[NTILE]:
//   X_ab += 2 sum_cd v P_cd, X_cd += 2 sum_ab v P_ab,
//   Y_ac += sum v P_bd, Y_ad += sum v P_bc, Y_bc += sum v P_ad, Y_bd += sum v P_ac,
// so that over all unique quartets J = X + X^T and K = Y + Y^T reproduce
// einsum('pqrs,rs') and einsum('prqs,rs') of hf.py:300.
template <int LA, int LB, int LC, int LD, typename VT>
SF_HD SF_INLINE void digest_quartet(const VT* __restrict__ v, const double* __restrict__ P,
                                    const QuartetMap& qm, double* __restrict__ t)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    double Pab[S::NA][S::NB], Pcd[S::NC][S::ND], Pac[S::NA][S::NC], Pad[S::NA][S::ND],
        Pbc[S::NB][S::NC], Pbd[S::NB][S::ND];
    const long long ld = qm.ld;
#pragma unroll
    for (int a = 0; a < S::NA; ++a) {
        const double* Pa = P + qm.row[0] + a * ld;
#pragma unroll
        for (int b = 0; b < S::NB; ++b) Pab[a][b] = Pa[qm.col[1] + b];
#pragma unroll
        for (int c = 0; c < S::NC; ++c) Pac[a][c] = Pa[qm.col[2] + c];
#pragma unroll
        for (int d = 0; d < S::ND; ++d) Pad[a][d] = Pa[qm.col[3] + d];
    }
#pragma unroll
    for (int b = 0; b < S::NB; ++b) {
        const double* Pb = P + qm.row[1] + b * ld;
#pragma unroll
        for (int c = 0; c < S::NC; ++c) Pbc[b][c] = Pb[qm.col[2] + c];
#pragma unroll
        for (int d = 0; d < S::ND; ++d) Pbd[b][d] = Pb[qm.col[3] + d];
    }
#pragma unroll
    for (int c = 0; c < S::NC; ++c)
#pragma unroll
        for (int d = 0; d < S::ND; ++d) Pcd[c][d] = P[qm.row[2] + c * ld + qm.col[3] + d];
#pragma unroll
    for (int i = 0; i < S::NTILE; ++i) t[i] = 0.0;
#pragma unroll
    for (int a = 0; a < S::NA; ++a)
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int c = 0; c < S::NC; ++c)
#pragma unroll
                for (int d = 0; d < S::ND; ++d) {
                    const double x = (double)v[((a * S::NB + b) * S::NC + c) * S::ND + d];
                    t[S::T_AB + a * S::NB + b] += 2.0 * x * Pcd[c][d];
                    t[S::T_CD + c * S::ND + d] += 2.0 * x * Pab[a][b];
                    t[S::T_AC + a * S::NC + c] += x * Pbd[b][d];
                    t[S::T_AD + a * S::ND + d] += x * Pbc[b][c];
                    t[S::T_BC + b * S::NC + c] += x * Pad[a][d];
                    t[S::T_BD + b * S::ND + d] += x * Pac[a][c];
                }
}

}  // namespace sf
