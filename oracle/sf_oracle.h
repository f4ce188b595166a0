/*
 * sf_oracle.h -- CPU FP64 oracle for the ERI / J-K path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (segfault_py_amd/ and the
 * HIP library) links, loads or calls this code.  It is used by tests/, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 *
 * It restates, in plain C, the reference's algorithms for this path:
 *   - rys.eri_rys / eri_3d_rys / ix / recur / shift / binom
 *       (src/hf/core/core.f90:222-229, 815-1002); the 5 Rys roots of root5
 *       (core.f90:232-812) are NOT copied: they come from a discretized
 *       Stieltjes procedure + Jacobi eigen-decomposition (Golub-Welsch).
 *   - mmd.boys / hermite_gauss_e / coulomb_auxiliary_r / eri_3d / eri
 *       (core.f90:23-198); Boys by series + downward recursion / erf + upward
 *       recursion instead of GSL 1F1.
 *   - one-electron S, T, V of mmd2.py:14-84.
 *   - build_eri_parallel's unique-quartet set and 8-fold scatter (hf.py:33-61).
 *
 * Basis functions are passed in the reference's per-Cartesian-component form
 * (ContractedGaussianFunction, molecule.py:201-226): for function f,
 * primitives [poff[f], poff[f]+npgf[f]) of the flat arrays coefs/expns/norms,
 * shell[3f..3f+2] = (lx,ly,lz), center[3f..3f+2] in bohr.
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* F_m(T), m = 0..mmax (mmax <= 32). */
void orc_boys(int mmax, double T, double *F);

/* Rys quadrature for weight exp(-X t^2) on t in [0,1]: nroots (1..8) nodes
 * u_r = t_r^2 (ascending) and weights w_r with
 * sum_r w_r u_r^k = F_k(X) for k = 0..2*nroots-1.  Returns 0 on success. */
int orc_rys_roots(int nroots, double X, double *u, double *w);

/* Five-node quadrature from fitted tables (the cost class of rys.root5,
 * core.f90:232-812): 16-term Chebyshev interpolants of orc_rys_roots on 128
 * X intervals of width 0.5, the half-range Gauss-Hermite limit for X >= 64.
 * orc_rys5_fast_init() builds the tables (not thread-safe; the whole-basis
 * functions below call it before their parallel loops). */
void orc_rys5_fast_init(void);
void orc_rys5_fast(double X, double *u, double *w);
/* Seconds one thread needs for n orc_rys5_fast calls, X spread over [0, xmax). */
double orc_time_rys5_fast(long n, double xmax);

/* One contracted Cartesian-component quartet, argument meaning of
 * rys.eri_rys (core.f90:967-1002): for each of the 4 functions, npgf[k]
 * primitives; the flat arrays hold the 4 functions' primitives back to back.
 * shells[12], centers[12]. */
double orc_eri_rys(const int npgf[4], const double *coefs, const double *expns,
                   const double *norms, const int shells[12], const double centers[12]);

/* Same quartet by McMurchie-Davidson (mmd.eri, core.f90:167-198). */
double orc_eri_mmd(const int npgf[4], const double *coefs, const double *expns,
                   const double *norms, const int shells[12], const double centers[12]);

/* Primitive building blocks, exposed for unit tests. */
double orc_hermite_e(int i, int j, int t, double qx, double a, double b);
double orc_coulomb_r(int t, int u, int v, int n, double p,
                     double pcx, double pcy, double pcz, double rpc);
double orc_eri_3d_rys(const int l[12], const double xyz[12], const double ex[4]);

/* Whole-basis operations over N functions (layout described above).
 * root_method: 0 = exact quadrature (orc_rys_roots), 1 = fitted tables
 * (orc_rys5_fast). */
void orc_eri_tensor(int N, const int *npgf, const int *poff, const double *coefs,
                    const double *expns, const double *norms, const int *shell,
                    const double *center, double *eri, int nthreads, int root_method);
void orc_eri_list(int N, const int *npgf, const int *poff, const double *coefs,
                  const double *expns, const double *norms, const int *shell,
                  const double *center, long nq, const int *quartets,
                  double *out, int nthreads, int root_method);
void orc_one_electron(int N, const int *npgf, const int *poff, const double *coefs,
                      const double *expns, const double *norms, const int *shell,
                      const double *center, int natom, const double *Z,
                      const double *xyz, double *S, double *T, double *V,
                      int nthreads);

#ifdef __cplusplus
}
#endif
#endif
