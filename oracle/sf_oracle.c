/*
 * sf_oracle.c -- CPU FP64 oracle for the ERI / J-K path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
 * benchmark's CPU baseline.  The product never links or calls it.
 *
 * Each routine cites the reference routine it restates (paths relative to
 * the reference repository root).  Parity of each function is pinned by the
 * tests named in its comment.
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include "sf_oracle.h"

#include <math.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338

/* ------------------------------------------------------------------------ */
/* Boys function.                                                           */
/* Reference: mmd.boys, src/hf/core/core.f90:23-28 (GSL 1F1(n+1/2;n+3/2;-T)  */
/* /(2n+1)) and utils.boys, src/hf/utils.py:107-109 (scipy hyp1f1).          */
/* Method (own): for T < 30 the series                                      */
/*   F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))              */
/* at m = mmax, then the downward recursion                                  */
/*   F_{m-1} = (2T F_m + e^{-T}) / (2m-1);                                  */
/* for T >= 30, F_0 = sqrt(pi/T) erf(sqrt T) / 2 and the upward recursion    */
/*   F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T).                                */
/* Pinned by tests/test_oracle.py (test_boys_*): mpmath hyp1f1 at 40 digits,*/
/* F_m(0) = 1/(2m+1), the F_0 erf closed form and the large-T asymptote.    */
/* ------------------------------------------------------------------------ */
void orc_boys(int mmax, double T, double *F)
{
    if (T < 0.0) T = 0.0;
    if (T < 30.0) {
        const double e = exp(-T);
        double term = 1.0 / (2.0 * mmax + 1.0);
        double sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * T / (2.0 * mmax + 2.0 * k + 1.0);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        F[mmax] = e * sum;
        for (int m = mmax; m > 0; --m)
            F[m - 1] = (2.0 * T * F[m] + e) / (2.0 * m - 1.0);
    } else {
        const double e = exp(-T);
        const double st = sqrt(T);
        F[0] = 0.5 * sqrt(ORC_PI / T) * erf(st);
        for (int m = 0; m < mmax; ++m)
            F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) / (2.0 * T);
    }
}

/* ------------------------------------------------------------------------ */
/* Rys roots and weights.                                                   */
/* Reference: rys.root5, src/hf/core/core.f90:232-812, which returns the 5   */
/* nodes as t^2 = RT/(1+RT) and their weights (core.f90:810-811).           */
/* The reference's piecewise fit tables are not used.  Method (own):        */
/*  1. discretize the measure exp(-X t^2) dt on [0, Te] (Te = 1, or          */
/*     sqrt(70/X) when X > 70 so the neglected tail is < e^-70 relative)    */
/*     with a 96-point Gauss-Legendre rule, as points u_k = t_k^2;           */
/*  2. discretized Stieltjes procedure (Gautschi 1982) for the recurrence    */
/*     coefficients alpha_j, beta_j of the monic orthogonal polynomials in u; */
/*  3. Golub-Welsch: eigen-decomposition of the Jacobi matrix by cyclic      */
/*     Jacobi rotations; node = eigenvalue, weight = beta_0 * v_0^2.          */
/* Pinned by tests/test_oracle.py (test_rys_*): mpmath Golub-Welsch, 50 dig.,*/
/* the moment identities sum w u^k = F_k(X), k < 2n.                         */
/* ------------------------------------------------------------------------ */
#define ORC_NGL 96
static double gl_x[ORC_NGL], gl_w[ORC_NGL];
static int gl_ready = 0;

static void gauss_legendre_init(void)
{
    /* Newton iteration on P_n with the three-term recurrence. */
    const int n = ORC_NGL;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(ORC_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) {
                /* one more evaluation at the converged point for pp */
                p1 = 1.0; p2 = 0.0;
                for (int j = 1; j <= n; ++j) {
                    double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
                }
                pp = n * (z * p1 - p2) / (z * z - 1.0);
                break;
            }
        }
        gl_x[i] = -z;
        gl_x[n - 1 - i] = z;
        gl_w[i] = gl_w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

static void jacobi_eigen(int n, double *A /* n*n, destroyed */, double *V /* n*n */)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off == 0.0) break;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                A[p * n + q] = A[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
}

int orc_rys_roots(int nroots, double X, double *u, double *w)
{
    if (nroots < 1 || nroots > 8) return 1;
    if (!gl_ready) {
        gauss_legendre_init();
        gl_ready = 1;
    }
    if (X < 0.0) X = 0.0;
    const double Te = (X > 70.0) ? sqrt(70.0 / X) : 1.0;
    double xk[ORC_NGL], wk[ORC_NGL], p0[ORC_NGL], p1[ORC_NGL];
    for (int k = 0; k < ORC_NGL; ++k) {
        double t = 0.5 * Te * (gl_x[k] + 1.0);
        xk[k] = t * t;
        wk[k] = 0.5 * Te * gl_w[k] * exp(-X * t * t);
        p0[k] = 0.0;
        p1[k] = 1.0;
    }
    double alpha[8], beta[8];
    double nrm_prev = 1.0;
    for (int j = 0; j < nroots; ++j) {
        double nrm = 0.0, xn = 0.0;
        for (int k = 0; k < ORC_NGL; ++k) {
            double q = wk[k] * p1[k] * p1[k];
            nrm += q;
            xn += q * xk[k];
        }
        alpha[j] = xn / nrm;
        beta[j] = (j == 0) ? nrm : nrm / nrm_prev;
        nrm_prev = nrm;
        for (int k = 0; k < ORC_NGL; ++k) {
            double pn = (xk[k] - alpha[j]) * p1[k] - ((j == 0) ? 0.0 : beta[j]) * p0[k];
            p0[k] = p1[k];
            p1[k] = pn;
        }
    }
    double A[64], V[64];
    const int n = nroots;
    memset(A, 0, sizeof(A));
    for (int i = 0; i < n; ++i) {
        A[i * n + i] = alpha[i];
        if (i + 1 < n) A[i * n + i + 1] = A[(i + 1) * n + i] = sqrt(beta[i + 1]);
    }
    jacobi_eigen(n, A, V);
    /* sort ascending */
    int idx[8];
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (A[idx[j] * n + idx[j]] < A[idx[i] * n + idx[i]]) {
                int t = idx[i]; idx[i] = idx[j]; idx[j] = t;
            }
    for (int i = 0; i < n; ++i) {
        int c = idx[i];
        u[i] = A[c * n + c];
        w[i] = beta[0] * V[0 * n + c] * V[0 * n + c];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Five-node Rys quadrature at the cost of a fitted evaluator.               */
/* Reference: rys.root5 (core.f90:232-812) evaluates piecewise polynomial    */
/* fits of degree <= 15 on X intervals and an asymptotic form for large X.   */
/* Method (own, with the same cost class): on X in [0, 64) cut into 128      */
/* intervals of width 0.5, each of the 5 nodes and 5 weights is a 16-term    */
/* Chebyshev interpolant of the exact quadrature above (orc_rys_roots,       */
/* sampled at the 16 Chebyshev points of the interval when the tables are    */
/* built); for X >= 64 the quadrature is that of the half-range Gaussian     */
/* exp(-X t^2) on [0, inf): u_r = h_r / X, w_r = g_r / sqrt(X) with          */
/* (h_r, g_r) = (X0 u_r(X0), sqrt(X0) w_r(X0)) from the exact solver at      */
/* X0 = 1e6 (the [1, inf) tail changes the moments by ~e^-X).               */
/* Used by orc_eri_* with root_method 1 (the CPU-baseline timing and large   */
/* test tensors); pinned by tests/test_oracle.py (test_rys5_fast_*) against */
/* the exact solver and the 50-digit mpmath vectors.                         */
/* ------------------------------------------------------------------------ */
#define F5_NI 128
#define F5_W 0.5
#define F5_XMAX (F5_NI * F5_W)
#define F5_M 16
static double f5_c[F5_NI][10][F5_M];
static double f5_h[5], f5_g[5];
static int f5_ready = 0;

void orc_rys5_fast_init(void)
{
    if (f5_ready) return;
    double u[5], w[5];
    for (int k = 0; k < F5_NI; ++k) {
        const double a = k * F5_W, b = a + F5_W;
        double f[F5_M][10];
        for (int j = 0; j < F5_M; ++j) {
            const double x = cos(ORC_PI * (j + 0.5) / F5_M);
            orc_rys_roots(5, 0.5 * (a + b) + 0.5 * (b - a) * x, u, w);
            for (int r = 0; r < 5; ++r) {
                f[j][r] = u[r];
                f[j][5 + r] = w[r];
            }
        }
        for (int v = 0; v < 10; ++v)
            for (int n = 0; n < F5_M; ++n) {
                double c = 0.0;
                for (int j = 0; j < F5_M; ++j) c += f[j][v] * cos(ORC_PI * n * (j + 0.5) / F5_M);
                f5_c[k][v][n] = 2.0 * c / F5_M;
            }
    }
    const double X0 = 1e6;
    orc_rys_roots(5, X0, u, w);
    for (int r = 0; r < 5; ++r) {
        f5_h[r] = u[r] * X0;
        f5_g[r] = w[r] * sqrt(X0);
    }
    f5_ready = 1;
}

/* 5 nodes (t^2, ascending) and weights; orc_rys5_fast_init() must have run. */
void orc_rys5_fast(double X, double *u, double *w)
{
    if (X < 0.0) X = 0.0;
    if (X >= F5_XMAX) {
        const double ix = 1.0 / X, isx = sqrt(ix);
        for (int r = 0; r < 5; ++r) {
            u[r] = f5_h[r] * ix;
            w[r] = f5_g[r] * isx;
        }
        return;
    }
    int k = (int)(X / F5_W);
    if (k >= F5_NI) k = F5_NI - 1;
    const double a = k * F5_W;
    const double s = (2.0 * (X - a) - F5_W) / F5_W;
    for (int v = 0; v < 10; ++v) {
        const double *c = f5_c[k][v];
        double b1 = 0.0, b2 = 0.0;
        for (int n = F5_M - 1; n >= 1; --n) {
            const double t = 2.0 * s * b1 - b2 + c[n];
            b2 = b1;
            b1 = t;
        }
        const double y = s * b1 - b2 + 0.5 * c[0];
        if (v < 5) u[v] = y;
        else w[v - 5] = y;
    }
}

/* Seconds one thread spends on n fitted five-node evaluations at X spread
 * over [0, xmax) (the root share of the CPU-baseline cost split). */
double orc_time_rys5_fast(long n, double xmax)
{
    orc_rys5_fast_init();
    double u[5], w[5], sink = 0.0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (long i = 0; i < n; ++i) {
        orc_rys5_fast(xmax * (double)((i * 7919) % n) / (double)n, u, w);
        sink += u[4] + w[0];
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    volatile double keep = sink;
    (void)keep;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------------ */
/* Reference Rys ERI path, restated.                                        */
/* ------------------------------------------------------------------------ */

/* rys.binom / fact, core.f90:213-229: C(n,k) from exact small factorials. */
static double orc_binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/* rys.recur, core.f90:815-874: 2-D Rys recurrence G(n,m), n <= li+lj,
 * m <= lk+ll, at node t (= t^2 of the quadrature). G is row-major with
 * (lk+ll+1) columns. */
static void orc_recur(int li, int lj, int lk, int ll, double xi, double xj, double xk,
                      double xl, double alpha, double beta, double gamma, double delta,
                      double t, double *G)
{
    const double a_ = alpha + beta, b_ = gamma + delta;
    const double xa = (alpha * xi + beta * xj) / a_;
    const double xb = (gamma * xk + delta * xl) / b_;
    const double ix0 = ORC_PI / sqrt(a_ * b_) *
                       exp(-alpha * beta / a_ * (xi - xj) * (xi - xj) -
                           gamma * delta / b_ * (xk - xl) * (xk - xl));
    const double b00 = 0.5 / (a_ + b_) * t;
    const double b10 = 0.5 / a_ - 0.5 * b_ / a_ / (a_ + b_) * t;
    const double b10p = 0.5 / b_ - 0.5 * a_ / b_ / (a_ + b_) * t;
    const double c00 = (xa - xi) + b_ / (a_ + b_) * (xb - xa) * t;
    const double c00p = (xb - xk) + a_ / (a_ + b_) * (xa - xb) * t;
    const int nn = li + lj, mm = lk + ll, ld = mm + 1;
#define G_(n, m) G[(n) * ld + (m)]
    for (int n = 0; n <= nn; ++n)
        for (int m = 0; m <= mm; ++m) G_(n, m) = 0.0;
    G_(0, 0) = ix0;
    if (nn > 0) G_(1, 0) = c00 * G_(0, 0);
    if (mm > 0) G_(0, 1) = c00p * G_(0, 0);
    for (int n = 1; n <= nn - 1; ++n) G_(n + 1, 0) = n * b10 * G_(n - 1, 0) + c00 * G_(n, 0);
    for (int m = 1; m <= mm - 1; ++m) G_(0, m + 1) = m * b10p * G_(0, m - 1) + c00p * G_(0, m);
    if (nn == 0 || mm == 0) return;
    for (int n = 1; n <= nn; ++n) {
        G_(n, 1) = n * b00 * G_(n - 1, 0) + c00p * G_(n, 0);
        for (int m = 2; m <= mm; ++m)
            G_(n, m) = b10p * (m - 1) * G_(n, m - 2) + n * b00 * G_(n - 1, m - 1) + c00p * G_(n, m - 1);
    }
#undef G_
}

/* rys.shift, core.f90:877-899: horizontal transfer on the 2-D integrals. */
static double orc_shift(const double *G, int ld, int li, int lj, int lk, int ll, double xij,
                        double xkl)
{
    double res = 0.0;
    for (int m = 0; m <= ll; ++m) {
        double tmp = 0.0;
        for (int n = 0; n <= lj; ++n)
            tmp += orc_binom(lj, n) * pow(xij, lj - n) * G[(n + li) * ld + (m + lk)];
        res += orc_binom(ll, m) * pow(xkl, ll - m) * tmp;
    }
    return res;
}

/* rys.ix, core.f90:902-914. */
static double orc_ix(int l1, int l2, int l3, int l4, double x1, double x2, double x3, double x4,
                     double a, double b, double c, double d, double t)
{
    double G[(8 + 1) * (8 + 1)];
    orc_recur(l1, l2, l3, l4, x1, x2, x3, x4, a, b, c, d, t, G);
    return orc_shift(G, l3 + l4 + 1, l1, l2, l3, l4, x1 - x2, x3 - x4);
}

/* rys.eri_3d_rys, core.f90:917-964: primitive quartet with 5 nodes, as
 * the reference (core.f90:952), up to total angular momentum L = 9, where 5
 * nodes integrate the polynomial part exactly.  Above (quartets with f
 * shells, L = 10..12) the reference's 5-node sum is no longer exact; the
 * oracle then takes floor(L/2) + 1 nodes of the exact quadrature, i.e. it
 * departs from the reference there on purpose and stays the exact integral
 * (pinned against the McMurchie-Davidson restatement, which has no
 * quadrature, in tests/test_oracle.py).
 * root_method 0 = the exact quadrature (orc_rys_roots), 1 = the fitted
 * 5-node evaluator (orc_rys5_fast, tables built) where 5 nodes are used. */
static double eri_3d_rys_m(const int l[12], const double xyz[12], const double ex[4],
                           int root_method)
{
    int L = 0;
    for (int k = 0; k < 12; ++k) L += l[k];
    const int nr = L / 2 + 1 > 5 ? L / 2 + 1 : 5;
    const double a = ex[0], b = ex[1], c = ex[2], d = ex[3];
    const double a_ = a + b, b_ = c + d, rho = a_ * b_ / (a_ + b_);
    double rpq2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double P = (a * xyz[0 + k] + b * xyz[3 + k]) / a_;
        double Q = (c * xyz[6 + k] + d * xyz[9 + k]) / b_;
        rpq2 += (P - Q) * (P - Q);
    }
    double u[8], w[8];
    if (root_method == 1 && nr == 5) orc_rys5_fast(rpq2 * rho, u, w);
    else orc_rys_roots(nr, rpq2 * rho, u, w);
    double res = 0.0;
    for (int r = 0; r < nr; ++r) {
        double ix = orc_ix(l[0], l[3], l[6], l[9], xyz[0], xyz[3], xyz[6], xyz[9], a, b, c, d, u[r]);
        double iy = orc_ix(l[1], l[4], l[7], l[10], xyz[1], xyz[4], xyz[7], xyz[10], a, b, c, d, u[r]);
        double iz = orc_ix(l[2], l[5], l[8], l[11], xyz[2], xyz[5], xyz[8], xyz[11], a, b, c, d, u[r]);
        res += ix * iy * iz * w[r];
    }
    return res * 2.0 * sqrt(rho / ORC_PI);
}

double orc_eri_3d_rys(const int l[12], const double xyz[12], const double ex[4])
{
    return eri_3d_rys_m(l, xyz, ex, 0);
}

/* rys.eri_rys, core.f90:967-1002: contraction over the 4 primitive sets. */
static double eri_rys_m(const int npgf[4], const double *coefs, const double *expns,
                        const double *norms, const int shells[12], const double centers[12],
                        int root_method)
{
    const int o1 = 0, o2 = npgf[0], o3 = o2 + npgf[1], o4 = o3 + npgf[2];
    int l[12];
    double xyz[12];
    for (int f = 0; f < 4; ++f)
        for (int k = 0; k < 3; ++k) {
            l[3 * f + k] = shells[3 * f + k];
            xyz[3 * f + k] = centers[3 * f + k];
        }
    double res = 0.0;
    for (int i1 = 0; i1 < npgf[0]; ++i1)
        for (int i2 = 0; i2 < npgf[1]; ++i2)
            for (int i3 = 0; i3 < npgf[2]; ++i3)
                for (int i4 = 0; i4 < npgf[3]; ++i4) {
                    double ex[4] = {expns[o1 + i1], expns[o2 + i2], expns[o3 + i3], expns[o4 + i4]};
                    double tmp = eri_3d_rys_m(l, xyz, ex, root_method);
                    res += norms[o1 + i1] * norms[o2 + i2] * norms[o3 + i3] * norms[o4 + i4] *
                           coefs[o1 + i1] * coefs[o2 + i2] * coefs[o3 + i3] * coefs[o4 + i4] * tmp;
                }
    return res;
}

double orc_eri_rys(const int npgf[4], const double *coefs, const double *expns,
                   const double *norms, const int shells[12], const double centers[12])
{
    return eri_rys_m(npgf, coefs, expns, norms, shells, centers, 0);
}

/* ------------------------------------------------------------------------ */
/* Reference McMurchie-Davidson path, restated (core.f90:30-198).           */
/* ------------------------------------------------------------------------ */

/* mmd.hermite_gauss_e, core.f90:30-61. */
double orc_hermite_e(int i, int j, int t, double qx, double a, double b)
{
    const double p = a + b, q = a * b / p;
    if (t < 0 || t > i + j) return 0.0;
    if (i == 0 && j == 0 && t == 0) return exp(-q * qx * qx);
    if (j == 0)
        return (0.5 / p) * orc_hermite_e(i - 1, j, t - 1, qx, a, b) -
               (q * qx / a) * orc_hermite_e(i - 1, j, t, qx, a, b) +
               (t + 1) * orc_hermite_e(i - 1, j, t + 1, qx, a, b);
    return (0.5 / p) * orc_hermite_e(i, j - 1, t - 1, qx, a, b) +
           (q * qx / b) * orc_hermite_e(i, j - 1, t, qx, a, b) +
           (t + 1) * orc_hermite_e(i, j - 1, t + 1, qx, a, b);
}

/* mmd.coulomb_auxiliary_r, core.f90:64-97. */
double orc_coulomb_r(int t, int u, int v, int n, double p, double pcx, double pcy, double pcz,
                     double rpc)
{
    double out = 0.0;
    if (t == 0 && u == 0 && v == 0) {
        double F[33];
        orc_boys(n, p * rpc * rpc, F);
        out = pow(-2.0 * p, n) * F[n];
    } else if (t == 0 && u == 0) {
        if (v >= 2) out += (v - 1) * orc_coulomb_r(t, u, v - 2, n + 1, p, pcx, pcy, pcz, rpc);
        out += pcz * orc_coulomb_r(t, u, v - 1, n + 1, p, pcx, pcy, pcz, rpc);
    } else if (t == 0) {
        if (u >= 2) out += (u - 1) * orc_coulomb_r(t, u - 2, v, n + 1, p, pcx, pcy, pcz, rpc);
        out += pcy * orc_coulomb_r(t, u - 1, v, n + 1, p, pcx, pcy, pcz, rpc);
    } else {
        if (t >= 2) out += (t - 1) * orc_coulomb_r(t - 2, u, v, n + 1, p, pcx, pcy, pcz, rpc);
        out += pcx * orc_coulomb_r(t - 1, u, v, n + 1, p, pcx, pcy, pcz, rpc);
    }
    return out;
}

/* mmd.eri_3d, core.f90:100-165 (E tables hoisted out of the 6-deep loop). */
static double orc_eri_3d_mmd(const int *s1, const double *c1, double a, const int *s2,
                             const double *c2, double b, const int *s3, const double *c3, double c,
                             const int *s4, const double *c4, double d)
{
    const double p = a + b, q = c + d, pq = p * q / (p + q);
    double P[3], Q[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = (a * c1[k] + b * c2[k]) / p;
        Q[k] = (c * c3[k] + d * c4[k]) / q;
    }
    const double rpq = sqrt((P[0] - Q[0]) * (P[0] - Q[0]) + (P[1] - Q[1]) * (P[1] - Q[1]) +
                            (P[2] - Q[2]) * (P[2] - Q[2]));
    double Eab[3][9], Ecd[3][9];
    for (int k = 0; k < 3; ++k) {
        for (int t = 0; t <= s1[k] + s2[k]; ++t)
            Eab[k][t] = orc_hermite_e(s1[k], s2[k], t, c1[k] - c2[k], a, b);
        for (int t = 0; t <= s3[k] + s4[k]; ++t)
            Ecd[k][t] = orc_hermite_e(s3[k], s4[k], t, c3[k] - c4[k], c, d);
    }
    double out = 0.0;
    for (int t = 0; t <= s1[0] + s2[0]; ++t)
        for (int u = 0; u <= s1[1] + s2[1]; ++u)
            for (int v = 0; v <= s1[2] + s2[2]; ++v)
                for (int tau = 0; tau <= s3[0] + s4[0]; ++tau)
                    for (int nu = 0; nu <= s3[1] + s4[1]; ++nu)
                        for (int phi = 0; phi <= s3[2] + s4[2]; ++phi) {
                            double sgn = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                            out += Eab[0][t] * Eab[1][u] * Eab[2][v] * Ecd[0][tau] * Ecd[1][nu] *
                                   Ecd[2][phi] * sgn *
                                   orc_coulomb_r(t + tau, u + nu, v + phi, 0, pq, P[0] - Q[0],
                                                 P[1] - Q[1], P[2] - Q[2], rpq);
                        }
    return out * 2.0 * pow(ORC_PI, 2.5) / (p * q * sqrt(p + q));
}

/* mmd.eri, core.f90:167-198 (the OpenMP race of the reference is absent:
 * this loop is serial). */
double orc_eri_mmd(const int npgf[4], const double *coefs, const double *expns,
                   const double *norms, const int shells[12], const double centers[12])
{
    const int o[4] = {0, npgf[0], npgf[0] + npgf[1], npgf[0] + npgf[1] + npgf[2]};
    double res = 0.0;
    for (int i1 = 0; i1 < npgf[0]; ++i1)
        for (int i2 = 0; i2 < npgf[1]; ++i2)
            for (int i3 = 0; i3 < npgf[2]; ++i3)
                for (int i4 = 0; i4 < npgf[3]; ++i4) {
                    double tmp = orc_eri_3d_mmd(shells + 0, centers + 0, expns[o[0] + i1],
                                                shells + 3, centers + 3, expns[o[1] + i2],
                                                shells + 6, centers + 6, expns[o[2] + i3],
                                                shells + 9, centers + 9, expns[o[3] + i4]);
                    res += norms[o[0] + i1] * norms[o[1] + i2] * norms[o[2] + i3] *
                           norms[o[3] + i4] * coefs[o[0] + i1] * coefs[o[1] + i2] *
                           coefs[o[2] + i3] * coefs[o[3] + i4] * tmp;
                }
    return res;
}

/* ------------------------------------------------------------------------ */
/* Whole-basis helpers.                                                     */
/* ------------------------------------------------------------------------ */
static double quartet_rys(int i, int j, int k, int l, const int *npgf, const int *poff,
                          const double *coefs, const double *expns, const double *norms,
                          const int *shell, const double *center, int root_method)
{
    const int f[4] = {i, j, k, l};
    int np[4], sh[12];
    double ce[12];
    double cf[4 * 64], ex[4 * 64], nr[4 * 64];
    int o = 0;
    for (int a = 0; a < 4; ++a) {
        np[a] = npgf[f[a]];
        for (int q = 0; q < np[a]; ++q) {
            cf[o + q] = coefs[poff[f[a]] + q];
            ex[o + q] = expns[poff[f[a]] + q];
            nr[o + q] = norms[poff[f[a]] + q];
        }
        o += np[a];
        for (int k3 = 0; k3 < 3; ++k3) {
            sh[3 * a + k3] = shell[3 * f[a] + k3];
            ce[3 * a + k3] = center[3 * f[a] + k3];
        }
    }
    return eri_rys_m(np, cf, ex, nr, sh, ce, root_method);
}

/* build_eri_parallel, src/hf/hf.py:33-61: unique quartets i>=j, k>=l,
 * ij >= kl with ij = i(i+1)/2 + j, each scattered to its 8 images. */
void orc_eri_tensor(int N, const int *npgf, const int *poff, const double *coefs,
                    const double *expns, const double *norms, const int *shell,
                    const double *center, double *eri, int nthreads, int root_method)
{
    const long n = N;
    if (root_method == 1) orc_rys5_fast_init();
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k < N; ++k)
                for (int l = 0; l <= k; ++l) {
                    if (i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
                    double v = quartet_rys(i, j, k, l, npgf, poff, coefs, expns, norms, shell, center,
                                           root_method);
                    const int P[8][4] = {{i, j, k, l}, {j, i, k, l}, {i, j, l, k}, {j, i, l, k},
                                         {k, l, i, j}, {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
                    for (int m = 0; m < 8; ++m)
                        eri[((P[m][0] * n + P[m][1]) * n + P[m][2]) * n + P[m][3]] = v;
                }
}

void orc_eri_list(int N, const int *npgf, const int *poff, const double *coefs,
                  const double *expns, const double *norms, const int *shell,
                  const double *center, long nq, const int *quartets, double *out, int nthreads,
                  int root_method)
{
    (void)N;
    if (root_method == 1) orc_rys5_fast_init();
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (long q = 0; q < nq; ++q)
        out[q] = quartet_rys(quartets[4 * q], quartets[4 * q + 1], quartets[4 * q + 2],
                             quartets[4 * q + 3], npgf, poff, coefs, expns, norms, shell, center,
                             root_method);
}

/* One-electron integrals, mmd2.py:14-84 (overlap_3d, kinetic_3d,
 * potential_3d and their contractions).  V includes -Z_C (hf.py:85-88). */
static double ovl3(double a, const int *s1, const double *c1, double b, const int *s2,
                   const double *c2)
{
    double sx = orc_hermite_e(s1[0], s2[0], 0, c1[0] - c2[0], a, b);
    double sy = orc_hermite_e(s1[1], s2[1], 0, c1[1] - c2[1], a, b);
    double sz = orc_hermite_e(s1[2], s2[2], 0, c1[2] - c2[2], a, b);
    return sx * sy * sz * pow(ORC_PI / (a + b), 1.5);
}

static double kin3(double a, const int *s1, const double *c1, double b, const int *s2,
                   const double *c2)
{
    const int l2 = s2[0], m2 = s2[1], n2 = s2[2];
    int t[3];
    double term0 = b * (2 * (l2 + m2 + n2) + 3) * ovl3(a, s1, c1, b, s2, c2);
    double term1 = 0.0, term2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        t[0] = l2; t[1] = m2; t[2] = n2;
        t[k] += 2;
        term1 += ovl3(a, s1, c1, b, t, c2);
        t[k] -= 4;
        if (s2[k] >= 2) term2 += s2[k] * (s2[k] - 1) * ovl3(a, s1, c1, b, t, c2);
    }
    return term0 - 2.0 * b * b * term1 - 0.5 * term2;
}

static double pot3(double a, const int *s1, const double *c1, double b, const int *s2,
                   const double *c2, const double *C)
{
    const double p = a + b;
    double P[3];
    for (int k = 0; k < 3; ++k) P[k] = (a * c1[k] + b * c2[k]) / p;
    double rpc = sqrt((P[0] - C[0]) * (P[0] - C[0]) + (P[1] - C[1]) * (P[1] - C[1]) +
                      (P[2] - C[2]) * (P[2] - C[2]));
    double res = 0.0;
    for (int t = 0; t <= s1[0] + s2[0]; ++t)
        for (int u = 0; u <= s1[1] + s2[1]; ++u)
            for (int v = 0; v <= s1[2] + s2[2]; ++v)
                res += orc_hermite_e(s1[0], s2[0], t, c1[0] - c2[0], a, b) *
                       orc_hermite_e(s1[1], s2[1], u, c1[1] - c2[1], a, b) *
                       orc_hermite_e(s1[2], s2[2], v, c1[2] - c2[2], a, b) *
                       orc_coulomb_r(t, u, v, 0, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], rpc);
    return res * 2.0 * ORC_PI / p;
}

void orc_one_electron(int N, const int *npgf, const int *poff, const double *coefs,
                      const double *expns, const double *norms, const int *shell,
                      const double *center, int natom, const double *Z, const double *xyz,
                      double *S, double *T, double *V, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0.0, t = 0.0, v = 0.0;
            for (int p = 0; p < npgf[i]; ++p)
                for (int q = 0; q < npgf[j]; ++q) {
                    const int ip = poff[i] + p, jq = poff[j] + q;
                    const double cc = norms[ip] * norms[jq] * coefs[ip] * coefs[jq];
                    const double a = expns[ip], b = expns[jq];
                    s += cc * ovl3(a, shell + 3 * i, center + 3 * i, b, shell + 3 * j, center + 3 * j);
                    t += cc * kin3(a, shell + 3 * i, center + 3 * i, b, shell + 3 * j, center + 3 * j);
                    for (int c = 0; c < natom; ++c)
                        v += -Z[c] * cc *
                             pot3(a, shell + 3 * i, center + 3 * i, b, shell + 3 * j,
                                  center + 3 * j, xyz + 3 * c);
                }
            S[(long)i * N + j] = s;
            T[(long)i * N + j] = t;
            V[(long)i * N + j] = v;
        }
}
