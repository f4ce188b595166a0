// sf_math.hpp -- Boys function and Rys quadrature for the ERI kernels.
//
// Written as host+device code: the HIP kernels use it on gfx950, and the
// CPU unit harness (tests) compiles the same header with g++ to check it.
//
// Reference counterparts: mmd.boys (src/hf/core/core.f90:23-28) and
// rys.root5 (core.f90:232-812).  The methods here are this project's own:
//  * Boys F_m(X): Taylor expansion about the nearest point of a 0.1-spaced
//    grid of exact values (X < 39.95) followed by downward recursion
//    F_{m-1} = (2X F_m + e^-X)/(2m-1); for X >= 39.95, F_0 = sqrt(pi/X)/2
//    (erfc(sqrt(40)) < 2e-19) and upward recursion.
//  * F_0, F_1 for one-node quadrature: alternating Maclaurin series for
//    X < 0.5, else F_0 = sqrt(pi) erf(sqrt X)/(2 sqrt X), F_1 = (F_0 - e^-X)/(2X).
//  * Rys nodes/weights for n = 1..5: unit-interval Chebyshev series of
//    degree 11 (tables generated at 50 digits by gen_tables.py) below
//    XMAX[n], the Gauss-Hermite large-X limit above.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SF_HD __host__ __device__
#define SF_DEV __device__
#define SF_INLINE __forceinline__
#else
#include <cmath>
#define SF_HD
#define SF_DEV
#define SF_INLINE inline
#endif

#include <math.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define SF_TABLE_QUAL static __device__ const
#else
#define SF_TABLE_QUAL static const
#endif
#include "rys_tables.inc"
#undef SF_TABLE_QUAL

namespace sf {

constexpr double kPi = 3.14159265358979323846264338;
constexpr double kSqrtPi = 1.77245385090551602729816748;

// ---------------------------------------------------------------- Boys ----
// 1/j! for the Taylor terms.
SF_HD SF_INLINE double inv_fact(int j)
{
    constexpr double t[10] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                              1.0 / 5040, 1.0 / 40320, 1.0 / 362880};
    return t[j];
}

// F_0..F_mmax at X >= 0 (mmax <= SF_BOYS_MTAB - 10).
SF_HD SF_INLINE void boys_grid(int mmax, double X, double* F)
{
    if (X < 39.95) {
        const int k = (int)(X * 10.0 + 0.5);
        const double d = k * SF_BOYS_DX - X;  // |d| <= 0.05
        const double* g = sf_boys_grid + k * SF_BOYS_MTAB + mmax;
        double s = g[8];
#pragma unroll
        for (int j = 7; j >= 0; --j) s = g[j] + s * d * (1.0 / (j + 1));
        // Horner form of sum_j g[j] d^j / j!
        F[mmax] = s;
        if (mmax > 0) {
            const double e = exp(-X);
            for (int m = mmax; m > 0; --m) F[m - 1] = (2.0 * X * F[m] + e) / (2.0 * m - 1.0);
        }
    } else {
        const double e = exp(-X);
        const double inv2x = 0.5 / X;
        F[0] = 0.5 * kSqrtPi / sqrt(X);
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) * inv2x;
    }
}

// F_0 and F_1 by closed form (no table loads).
SF_HD SF_INLINE void boys01(double X, double& F0, double& F1)
{
    if (X < 0.5) {
        // F_m(X) = sum_k (-X)^k / (k! (2m+2k+1)), k = 0..15
        double s0 = 0.0, s1 = 0.0, term = 1.0;
        double pw[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            pw[k] = term;
            term *= -X / (k + 1);
        }
#pragma unroll
        for (int k = 15; k >= 0; --k) {
            s0 += pw[k] / (2 * k + 1);
            s1 += pw[k] / (2 * k + 3);
        }
        F0 = s0;
        F1 = s1;
    } else {
        const double sx = sqrt(X);
        const double e = exp(-X);
        F0 = 0.5 * kSqrtPi * erf(sx) / sx;
        F1 = (F0 - e) * (0.5 / X);
    }
}

// ----------------------------------------------------------------- Rys ----
template <int N>
SF_HD SF_INLINE const double* rys_table()
{
    if constexpr (N == 1) return sf_rys_cheb1;
    else if constexpr (N == 2) return sf_rys_cheb2;
    else if constexpr (N == 3) return sf_rys_cheb3;
    else if constexpr (N == 4) return sf_rys_cheb4;
    else return sf_rys_cheb5;
}

template <int N>
SF_HD SF_INLINE double herm_u(int r)
{
    if constexpr (N == 1) return sf_herm_u1[r];
    else if constexpr (N == 2) return sf_herm_u2[r];
    else if constexpr (N == 3) return sf_herm_u3[r];
    else if constexpr (N == 4) return sf_herm_u4[r];
    else return sf_herm_u5[r];
}

template <int N>
SF_HD SF_INLINE double herm_w(int r)
{
    if constexpr (N == 1) return sf_herm_w1[r];
    else if constexpr (N == 2) return sf_herm_w2[r];
    else if constexpr (N == 3) return sf_herm_w3[r];
    else if constexpr (N == 4) return sf_herm_w4[r];
    else return sf_herm_w5[r];
}

constexpr int rys_xmax(int n) { return n == 1 ? 40 : n == 2 ? 48 : n == 3 ? 54 : n == 4 ? 60 : 66; }

// Nodes u_r = t_r^2 (ascending) and weights w_r of the N-point Rys rule.
template <int N>
SF_HD SF_INLINE void rys_roots(double X, double* u, double* w)
{
    static_assert(N >= 1 && N <= 5, "1..5 Rys nodes");
    if (X < (double)rys_xmax(N)) {
        const int it = (int)X;
        const double s = 2.0 * (X - it) - 1.0;
        const double s2 = 2.0 * s;
        const double* c = rys_table<N>() + it * (2 * N) * (SF_RYS_DEG + 1);
#pragma unroll
        for (int v = 0; v < 2 * N; ++v) {
            const double* cv = c + v * (SF_RYS_DEG + 1);
            // Clenshaw recurrence for sum_j c_j T_j(s)
            double b1 = 0.0, b2 = 0.0;
#pragma unroll
            for (int j = SF_RYS_DEG; j >= 1; --j) {
                const double b0 = cv[j] + s2 * b1 - b2;
                b2 = b1;
                b1 = b0;
            }
            const double val = cv[0] + s * b1 - b2;
            if (v < N) u[v] = val;
            else w[v - N] = val;
        }
    } else {
        const double ix = 1.0 / X;
        const double isx = sqrt(ix);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            u[r] = herm_u<N>(r) * ix;
            w[r] = herm_w<N>(r) * isx;
        }
    }
}

// One-node rule from F_0, F_1 (used where only one node is needed).
SF_HD SF_INLINE void rys_roots1_closed(double X, double* u, double* w)
{
    double F0, F1;
    boys01(X, F0, F1);
    u[0] = F1 / F0;
    w[0] = F0;
}

}  // namespace sf
