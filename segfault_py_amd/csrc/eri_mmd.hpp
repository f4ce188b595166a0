// eri_mmd.hpp -- McMurchie-Davidson shell-quartet ERIs with the Hermite ->
// Cartesian transforms (E . R . E) on the matrix cores (gfx950 MFMA, FP64).
//
// Reference counterpart: mmd.eri / mmd.eri_3d (src/hf/core/core.f90:100-198),
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
//             sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi}
//             R_{t+tau, u+nu, v+phi}(alpha, PQ),
// there a 6-deep scalar loop over recursive E and R per Cartesian component
// and primitive quartet.  Here one wavefront takes one contracted shell
// quartet, all components at once:
//   * per primitive quartet, the Boys values F_0..F_L(alpha |PQ|^2) and the
//     Hermite integrals R^0_{tuv} (t+u+v <= L) are built level by level in
//     LDS, then  T[h_ab][m_cd] += pref R[h_ab + h_cd] (-1)^{h_cd} E^{cd}[m_cd][h_cd]
//     is one MFMA GEMM (16x16x4 f64 tiles), accumulated over the ket
//     primitives in the matrix accumulators;
//   * per bra primitive,  Out[m_ab][m_cd] += E^{ab}[m_ab][h_ab] T[h_ab][m_cd]
//     is the second GEMM, T staged through LDS;
//   * Out is written as raw integral planes out[(m_ab*M_cd + m_cd)*qstride + q]
//     in the shell-normalised basis, the layout of the Rys raw planes.
// h indexes Hermite triples (t,u,v) ordered by t+u+v, so the triples of
// total <= S are the first (S+1)(S+2)(S+3)/6 indices.
#pragma once
#include "mmd.hpp"

namespace sf {

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

constexpr int herm_n(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Hermite triple of index h and the index of a triple, for t+u+v <= 8.
struct HermIdx {
    int t[165], u[165], v[165];
    int idx[9][9][9];
    constexpr HermIdx() : t(), u(), v(), idx()
    {
        int h = 0;
        for (int s = 0; s <= 8; ++s)
            for (int tt = s; tt >= 0; --tt)
                for (int uu = s - tt; uu >= 0; --uu) {
                    const int vv = s - tt - uu;
                    t[h] = tt;
                    u[h] = uu;
                    v[h] = vv;
                    idx[tt][uu][vv] = h;
                    ++h;
                }
    }
};
__device__ constexpr HermIdx kHerm{};

template <int LA, int LB, int LC, int LD>
struct MmdShape {
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int HAB = herm_n(LAB), HCD = herm_n(LCD), HL = herm_n(L);
    static constexpr int NA = cart_n(LA), NB = cart_n(LB), NC = cart_n(LC), ND = cart_n(LD);
    static constexpr int MAB = NA * NB, MCD = NC * ND;
    static constexpr int BI = (HAB + 15) / 16;  // T row blocks (h_ab)
    static constexpr int BJ = (MCD + 15) / 16;  // column blocks (m_cd), both GEMMs
    static constexpr int K1 = (HCD + 3) / 4;    // first GEMM k steps (h_cd)
    static constexpr int BM = (MAB + 15) / 16;  // Out row blocks (m_ab)
    static constexpr int K2 = (HAB + 3) / 4;    // second GEMM k steps (h_ab)
    static constexpr int TR = BI * 16, TC = BJ * 16;  // T staging tile in LDS
    static constexpr int EAB = (LA + 1) * (LB + 1) * (LAB + 1);
    static constexpr int ECD = (LC + 1) * (LD + 1) * (LCD + 1);
};

// One direction's E table of a primitive pair, flattened [i][j][t].
template <int I, int J>
__device__ __forceinline__ void store_E(double ip2, double PA, double PB, double* __restrict__ dst)
{
    double E[I + 1][J + 1][I + J + 1];
    hermite_E<I, J>(ip2, PA, PB, E);
#pragma unroll
    for (int i = 0; i <= I; ++i)
#pragma unroll
        for (int j = 0; j <= J; ++j)
#pragma unroll
            for (int t = 0; t <= I + J; ++t) dst[(i * (J + 1) + j) * (I + J + 1) + t] = E[i][j][t];
}

// E^{ab}[m][h] from the three direction tables (m = fa * NB + fb); zero
// outside the matrix.
template <int LA, int LB>
__device__ __forceinline__ double e_elem(const double (*__restrict__ E)[(LA + 1) * (LB + 1) * (LA + LB + 1)],
                                         int m, int h, int nh)
{
    constexpr int NB = cart_n(LB), NM = cart_n(LA) * NB, NT = LA + LB + 1;
    if (m >= NM || h >= nh) return 0.0;
    const int fa = m / NB, fb = m - fa * NB;
    const int t = kHerm.t[h], u = kHerm.u[h], v = kHerm.v[h];
    const int ax = cart_l(LA, fa, 0), ay = cart_l(LA, fa, 1), az = cart_l(LA, fa, 2);
    const int bx = cart_l(LB, fb, 0), by = cart_l(LB, fb, 1), bz = cart_l(LB, fb, 2);
    if (t > ax + bx || u > ay + by || v > az + bz) return 0.0;
    return E[0][(ax * (LB + 1) + bx) * NT + t] * E[1][(ay * (LB + 1) + by) * NT + u] *
           E[2][(az * (LB + 1) + bz) * NT + v];
}

// All component integrals of one shell quartet (bra pairs bra[0..nbra),
// ket pairs ket[0..nket)) into out[(m_ab * MCD + m_cd) * qstride + qa].
// Called by all 64 lanes of a one-wavefront block with wave-uniform
// arguments; uses the block's LDS arrays.
template <int LA, int LB, int LC, int LD>
__device__ void mmd_shell_quartet(const PrimPair* __restrict__ bra, int nbra,
                                  const PrimPair* __restrict__ ket, int nket, const double* A,
                                  const double* B, const double* C, const double* D,
                                  double* __restrict__ out, long long qstride, long long qa,
                                  double (*sR)[herm_n(LA + LB + LC + LD)],
                                  double (*sEab)[MmdShape<LA, LB, LC, LD>::EAB],
                                  double (*sEcd)[MmdShape<LA, LB, LC, LD>::ECD],
                                  double* __restrict__ sT)
{
    using M = MmdShape<LA, LB, LC, LD>;
    constexpr int L = M::L;
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;  // MFMA operand row/column and k slot
    mfma_d4 oacc[M::BM][M::BJ];
#pragma unroll
    for (int x = 0; x < M::BM; ++x)
#pragma unroll
        for (int y = 0; y < M::BJ; ++y) oacc[x][y] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    for (int ib = 0; ib < nbra; ++ib) {
        const PrimPair pb = bra[ib];
        const double p = pb.p;
        if (lane < 3) {
            const double P = lane == 0 ? pb.Px : lane == 1 ? pb.Py : pb.Pz;
            store_E<LA, LB>(0.5 * pb.ip, P - A[lane], P - B[lane], sEab[lane]);
        }
        mfma_d4 tacc[M::BI][M::BJ];
#pragma unroll
        for (int x = 0; x < M::BI; ++x)
#pragma unroll
            for (int y = 0; y < M::BJ; ++y) tacc[x][y] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        for (int ik = 0; ik < nket; ++ik) {
            const PrimPair pk = ket[ik];
            const double q = pk.p;
            if (lane >= 3 && lane < 6) {
                const int k = lane - 3;
                const double Qk = k == 0 ? pk.Px : k == 1 ? pk.Py : pk.Pz;
                store_E<LC, LD>(0.5 * pk.ip, Qk - C[k], Qk - D[k], sEcd[k]);
            }
            const double pq = p + q;
            const double rs = fast_rsqrt(pq);
            const double alpha = p * q * rs * rs;
            const double PQ[3] = {pb.Px - pk.Px, pb.Py - pk.Py, pb.Pz - pk.Pz};
            const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725 * pb.Kp * pk.Kp * rs;
            double F[L + 1];
            boys_grid(L, T, F);
            // R^n_{tuv}, n = L .. 0: level n holds t+u+v <= L-n in sR[n & 1]
            const double m2a = -2.0 * alpha;
            double s = 1.0;
#pragma unroll
            for (int n = 0; n < L; ++n) s *= m2a;  // (-2 alpha)^L
            if (lane == 0) sR[L & 1][0] = s * F[L];
            __syncthreads();
#pragma unroll
            for (int n = L - 1; n >= 0; --n) {
                s = (n == 0) ? 1.0 : s / m2a;  // (-2 alpha)^n
                const double* Rp = sR[(n + 1) & 1];
                double* Rn = sR[n & 1];
                for (int h = lane; h < herm_n(L - n); h += 64) {
                    const int t = kHerm.t[h], u = kHerm.u[h], v = kHerm.v[h];
                    double r;
                    if (t > 0) {
                        r = PQ[0] * Rp[kHerm.idx[t - 1][u][v]];
                        if (t > 1) r += (t - 1) * Rp[kHerm.idx[t - 2][u][v]];
                    } else if (u > 0) {
                        r = PQ[1] * Rp[kHerm.idx[t][u - 1][v]];
                        if (u > 1) r += (u - 1) * Rp[kHerm.idx[t][u - 2][v]];
                    } else if (v > 0) {
                        r = PQ[2] * Rp[kHerm.idx[t][u][v - 1]];
                        if (v > 1) r += (v - 1) * Rp[kHerm.idx[t][u][v - 2]];
                    } else {
                        r = s * F[n];
                    }
                    Rn[h] = r;
                }
                __syncthreads();
            }
            const double* R0 = sR[0];
            // T += pref R E_cd^T  (rows h_ab, columns m_cd, k = h_cd)
#pragma unroll 1
            for (int kk = 0; kk < M::K1; ++kk) {
                const int hc = kk * 4 + lk;
                double bop[M::BJ];
                const bool hv = hc < M::HCD;
                const int sg = hv ? (((kHerm.t[hc] + kHerm.u[hc] + kHerm.v[hc]) & 1) ? -1 : 1) : 0;
#pragma unroll
                for (int y = 0; y < M::BJ; ++y)
                    bop[y] = hv ? sg * e_elem<LC, LD>(sEcd, y * 16 + lr, hc, M::HCD) : 0.0;
#pragma unroll
                for (int x = 0; x < M::BI; ++x) {
                    const int ha = x * 16 + lr;
                    double aop = 0.0;
                    if (hv && ha < M::HAB)
                        aop = pref * R0[kHerm.idx[kHerm.t[ha] + kHerm.t[hc]][kHerm.u[ha] + kHerm.u[hc]]
                                                 [kHerm.v[ha] + kHerm.v[hc]]];
#pragma unroll
                    for (int y = 0; y < M::BJ; ++y)
                        tacc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bop[y], tacc[x][y], 0, 0, 0);
                }
            }
            __syncthreads();  // sR and sEcd are rewritten by the next ket primitive
        }
        // T -> LDS: accumulator element r of tile (x, y) is row x*16 + lk + 4r,
        // column y*16 + lr (v_mfma_f64_16x16x4f64 with A[lr][lk], B[lk][lr])
#pragma unroll
        for (int x = 0; x < M::BI; ++x)
#pragma unroll
            for (int y = 0; y < M::BJ; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) sT[(x * 16 + lk + 4 * r) * M::TC + y * 16 + lr] = tacc[x][y][r];
        __syncthreads();
        // Out += E_ab T  (rows m_ab, columns m_cd, k = h_ab)
#pragma unroll 1
        for (int kk = 0; kk < M::K2; ++kk) {
            const int ha = kk * 4 + lk;
            double bop[M::BJ];
#pragma unroll
            for (int y = 0; y < M::BJ; ++y) bop[y] = sT[ha * M::TC + y * 16 + lr];
#pragma unroll
            for (int x = 0; x < M::BM; ++x) {
                const double aop = e_elem<LA, LB>(sEab, x * 16 + lr, ha, M::HAB);
#pragma unroll
                for (int y = 0; y < M::BJ; ++y)
                    oacc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bop[y], oacc[x][y], 0, 0, 0);
            }
        }
        __syncthreads();  // sEab and sT are rewritten by the next bra primitive
    }
#pragma unroll
    for (int x = 0; x < M::BM; ++x)
#pragma unroll
        for (int y = 0; y < M::BJ; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int mab = x * 16 + lk + 4 * r, mcd = y * 16 + lr;
                if (mab < M::MAB && mcd < M::MCD)
                    out[(long long)(mab * M::MCD + mcd) * qstride + qa] = oacc[x][y][r];
            }
}

}  // namespace sf
