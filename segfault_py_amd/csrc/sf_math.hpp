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
//  * F_0, F_1 for one-node quadrature: the same grid Taylor series for both
//    (no exp) below X = 40, F_0 = sqrt(pi/X)/2, F_1 = F_0/(2X) above.
//  * Rys nodes/weights for n = 1..7: unit-interval Chebyshev series of
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

// ------------------------------------------------- reciprocal, 1/sqrt ----
// Hardware seeds (v_rcp_f64 / v_rsq_f64) refined by two Newton steps to
// full double precision, for the positive normal arguments of the kernels
// (p + q, X >= 39.95, F_0).  The host build seeds from the exact value with
// a deliberate 3e-9 relative error so the CPU harness exercises the
// refinement itself.
SF_HD SF_INLINE double fast_rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
#else
    double y = (1.0 / x) * (1.0 + 3e-9);
#endif
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

SF_HD SF_INLINE double fast_rsqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
#else
    double y = (1.0 / sqrt(x)) * (1.0 + 3e-9);
#endif
    double e = fma(-x * y, y, 1.0);  // 1 - x y^2
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// ---------------------------------------------------------------- Boys ----
// 1/(2m-1) for the downward recursion, m = 1..34.
SF_HD SF_INLINE double inv_odd(int m)
{
    constexpr double t[35] = {
        -1.0,       1.0,        1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 9,  1.0 / 11, 1.0 / 13,
        1.0 / 15,   1.0 / 17,   1.0 / 19, 1.0 / 21, 1.0 / 23, 1.0 / 25, 1.0 / 27, 1.0 / 29,
        1.0 / 31,   1.0 / 33,   1.0 / 35, 1.0 / 37, 1.0 / 39, 1.0 / 41, 1.0 / 43, 1.0 / 45,
        1.0 / 47,   1.0 / 49,   1.0 / 51, 1.0 / 53, 1.0 / 55, 1.0 / 57, 1.0 / 59, 1.0 / 61,
        1.0 / 63,   1.0 / 65,   1.0 / 67};
    return t[m];
}

// sum_{j=0}^{8} g[j] d^j / j!  (Horner; the 1/(j+1) factors are constants)
SF_HD SF_INLINE double taylor9(const double* g, double d)
{
    double s = g[8];
    s = g[7] + s * (d * (1.0 / 8));
    s = g[6] + s * (d * (1.0 / 7));
    s = g[5] + s * (d * (1.0 / 6));
    s = g[4] + s * (d * (1.0 / 5));
    s = g[3] + s * (d * (1.0 / 4));
    s = g[2] + s * (d * (1.0 / 3));
    s = g[1] + s * (d * (1.0 / 2));
    return g[0] + s * d;
}

// F_0..F_mmax at X >= 0 (mmax <= SF_BOYS_MTAB - 10).
SF_HD SF_INLINE void boys_grid(int mmax, double X, double* F)
{
    if (X < 39.95) {
        const int k = (int)(X * 10.0 + 0.5);
        const double d = k * SF_BOYS_DX - X;  // |d| <= 0.05
        F[mmax] = taylor9(sf_boys_grid + k * SF_BOYS_MTAB + mmax, d);
        if (mmax > 0) {
            const double e = exp(-X);
            for (int m = mmax; m > 0; --m) F[m - 1] = (2.0 * X * F[m] + e) * inv_odd(m);
        }
    } else {
        const double e = exp(-X);
        const double ix = 1.0 / X;
        F[0] = 0.5 * kSqrtPi * sqrt(ix);
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) * (0.5 * ix);
    }
}

// F_0 and F_1 without exp or erf: grid Taylor series below X = 40, the
// asymptotic forms above (e^-X < 4.3e-18 there, below double rounding of
// F_0 and F_1).  When only F_0 is used the compiler drops the F_1 terms.
// g: grid rows of F_0..F_9 at X_k = 0.1 k, `stride` doubles apart (the
// global table sf_boys_grid has stride SF_BOYS_MTAB; kernels may stage the
// first 10 columns in LDS with stride 10).
SF_HD SF_INLINE void boys01_tab(double X, const double* g, int stride, double& F0, double& F1)
{
    if (X < 39.95) {
        const int k = (int)(X * 10.0 + 0.5);
        const double d = k * SF_BOYS_DX - X;
        const double* gk = g + k * stride;
        F0 = taylor9(gk, d);
        F1 = taylor9(gk + 1, d);
    } else {
        F0 = 0.5 * kSqrtPi * fast_rsqrt(X);
        F1 = 0.5 * F0 * fast_rcp(X);
    }
}

SF_HD SF_INLINE void boys01(double X, double& F0, double& F1)
{
    boys01_tab(X, sf_boys_grid, SF_BOYS_MTAB, F0, F1);
}

// The one-node kernels' LDS table: rows at X_k = 0.2 k (k = 0..200, the even
// rows of the global grid) of 12 Taylor coefficients c_m = F_m(X_k) / m!.
// 19.3 KB instead of the 32 KB of 400 rows of F_0..F_9, so 8 blocks of 256
// threads fit a CU's LDS instead of 5.
// (A 0.1 grid of 9 coefficients -- 6 fewer FMAs per primitive quartet,
// 28.9 KB -- measured no faster on the 20-water build: 6.41-6.56 ms against
// 6.37-6.40 ms.)
constexpr int kBoysTaylorRows = 201;
constexpr int kBoysTaylorCols = 12;
constexpr double kBoysTaylorDX = 0.2;
// rows of the global 0.1 grid per Taylor row
constexpr int kBoysTaylorStep = 2;

// 1/m! for m <= 11.
SF_HD SF_INLINE double inv_fact(int m)
{
    constexpr double t[12] = {1.0,           1.0,            1.0 / 2,         1.0 / 6,
                              1.0 / 24,      1.0 / 120,      1.0 / 720,       1.0 / 5040,
                              1.0 / 40320,   1.0 / 362880,   1.0 / 3628800,   1.0 / 39916800};
    return t[m];
}

// Quartets whose every primitive quartet has X >= kFarX are "far": there
// boys01_taylor takes its asymptotic branch (X >= 39.95; the neglected
// erfc tail is below 1e-17 relative), so F_0 = sqrt(pi / (4X)) and
// F_1 = F_0 / (2X) exactly as in that branch.
constexpr double kFarX = 40.0;

// F_0 and F_1 from the Taylor rows: with d = X_k - X (|d| <= 0.1),
// F_0(X) = sum_m c_m d^m and F_1 = -dF_0/dX = dF_0/dd = sum_m m c_m d^(m-1),
// both from one Horner pass (truncation below 1e-19 for both).
SF_HD SF_INLINE void boys01_taylor(double X, const double* c, double& F0, double& F1)
{
    if (X < 39.95) {
        const int k = (int)(X * (1.0 / kBoysTaylorDX) + 0.5);
        const double d = k * kBoysTaylorDX - X;
        const double* ck = c + k * kBoysTaylorCols;
        double p = ck[kBoysTaylorCols - 1], dp = 0.0;
#pragma unroll
        for (int m = kBoysTaylorCols - 2; m >= 0; --m) {
            dp = dp * d + p;
            p = p * d + ck[m];
        }
        F0 = p;
        F1 = dp;
    } else {
        F0 = 0.5 * kSqrtPi * fast_rsqrt(X);
        F1 = 0.5 * F0 * fast_rcp(X);
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
    else if constexpr (N == 5) return sf_rys_cheb5;
    else if constexpr (N == 6) return sf_rys_cheb6;
    else return sf_rys_cheb7;
}

template <int N>
SF_HD SF_INLINE double herm_u(int r)
{
    if constexpr (N == 1) return sf_herm_u1[r];
    else if constexpr (N == 2) return sf_herm_u2[r];
    else if constexpr (N == 3) return sf_herm_u3[r];
    else if constexpr (N == 4) return sf_herm_u4[r];
    else if constexpr (N == 5) return sf_herm_u5[r];
    else if constexpr (N == 6) return sf_herm_u6[r];
    else return sf_herm_u7[r];
}

template <int N>
SF_HD SF_INLINE double herm_w(int r)
{
    if constexpr (N == 1) return sf_herm_w1[r];
    else if constexpr (N == 2) return sf_herm_w2[r];
    else if constexpr (N == 3) return sf_herm_w3[r];
    else if constexpr (N == 4) return sf_herm_w4[r];
    else if constexpr (N == 5) return sf_herm_w5[r];
    else if constexpr (N == 6) return sf_herm_w6[r];
    else return sf_herm_w7[r];
}

constexpr int rys_xmax(int n)
{
    return n == 1 ? 40 : n == 2 ? 48 : n == 3 ? 54 : n == 4 ? 60 : n == 5 ? 66 : n == 6 ? 72 : 78;
}

// Nodes u_r = t_r^2 (ascending) and weights w_r of the N-point Rys rule,
// from the Chebyshev table c of N nodes (rys_table<N>() or a copy of it).
template <int N>
SF_HD SF_INLINE void rys_roots_tab(double X, const double* __restrict__ tab, double* u, double* w)
{
    static_assert(N >= 1 && N <= SF_RYS_NMAX, "1..7 Rys nodes");
    if (X < (double)rys_xmax(N)) {
        const int it = (int)X;
        const double s = 2.0 * (X - it) - 1.0;
        const double s2 = 2.0 * s;
        const double* c = tab + it * (2 * N) * (SF_RYS_DEG + 1);
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
        const double ix = fast_rcp(X);
        const double isx = fast_rsqrt(X);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            u[r] = herm_u<N>(r) * ix;
            w[r] = herm_w<N>(r) * isx;
        }
    }
}

template <int N>
SF_HD SF_INLINE void rys_roots(double X, double* u, double* w)
{
    rys_roots_tab<N>(X, rys_table<N>(), u, w);
}

// One-node rule from F_0, F_1 (used where only one node is needed).
SF_HD SF_INLINE void rys_roots1_closed(double X, double* u, double* w)
{
    double F0, F1;
    boys01(X, F0, F1);
    u[0] = F1 * fast_rcp(F0);
    w[0] = F0;
}

// The node table a kernel reads: for one node the Boys grid rows (F_0..F_9,
// `stride` doubles apart), for N >= 2 nodes the Chebyshev table of N nodes.
struct NodeTable {
    const double* p;
    int stride;
    // one node: 0 = rows of F_0..F_9 (the global grid), 1 = the Taylor rows
    // of boys01_taylor (the kernels' LDS copy)
    int taylor = 0;
};

// F_0 and F_1 from a one-node table of either kind.
SF_HD SF_INLINE void boys01_node(double X, const NodeTable& t, double& F0, double& F1)
{
    if (t.taylor) boys01_taylor(X, t.p, F0, F1);
    else boys01_tab(X, t.p, t.stride, F0, F1);
}

template <int N>
SF_HD SF_INLINE NodeTable global_node_table()
{
    if constexpr (N == 1) return NodeTable{sf_boys_grid, SF_BOYS_MTAB};
    else return NodeTable{rys_table<N>(), 0};
}

// Doubles a kernel stages in LDS for N nodes (0: read the global table):
// the Taylor rows of boys01_taylor for N = 1, the Chebyshev table for
// N = 2, 3.
constexpr int node_table_lds(int n)
{
    return n == 1 ? kBoysTaylorRows * kBoysTaylorCols : n == 2 ? 48 * 4 * (SF_RYS_DEG + 1) : n == 3 ? 54 * 6 * (SF_RYS_DEG + 1) : 0;
}

// Element k of the LDS copy for N nodes.
template <int N>
SF_HD SF_INLINE double node_table_elem(int k)
{
    if constexpr (N == 1)
        return sf_boys_grid[kBoysTaylorStep * (k / kBoysTaylorCols) * SF_BOYS_MTAB + k % kBoysTaylorCols] *
               inv_fact(k % kBoysTaylorCols);
    else return rys_table<N>()[k];
}

template <int N>
SF_HD SF_INLINE void rys_nodes_tab(double X, const NodeTable& t, double* u, double* w)
{
    if constexpr (N == 1) {
        double F0, F1;
        boys01_node(X, t, F0, F1);
        u[0] = F1 * fast_rcp(F0);
        w[0] = F0;
    } else {
        rys_roots_tab<N>(X, t.p, u, w);
    }
}

}  // namespace sf
