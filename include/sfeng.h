/*
 * sfeng.h -- C-ABI of the MI355X (gfx950) ERI and J/K (Fock-build) engine.
 *
 * Conventions
 *   - Every entry point returns an int status: 0 = OK, 1 = bad argument,
 *     2 = HIP runtime error, 3 = out of memory, 4 = wrong call order,
 *     5 = unsupported size / angular momentum.  sf_last_error() gives the
 *     calling thread's last message.  No exception or abort crosses the ABI.
 *   - Buffers are caller-owned.  "device" pointers are HIP device memory
 *     (e.g. torch.cuda tensors' data_ptr()); "host" pointers are host memory.
 *   - `stream` is a hipStream_t (NULL = the default stream).
 *   - Functions are numbered (ij|kl) in chemists' notation, in the AO order
 *     of the reference (src/hf/molecule.py:138-165): atom by atom, shell by
 *     shell, Cartesian components s; x,y,z; xx,xy,xz,yy,yz,zz.
 *   - Each Cartesian component is normalised on its own
 *     (molecule.py:210-226); see sf_basis_create for how that is passed.
 */
#ifndef SFENG_H
#define SFENG_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_basis sf_basis;

/* Last error message of the calling thread (empty string if none). */
const char *sf_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int sf_version(void);

/* ---- Boys function and Rys quadrature (device buffers) ----------------
 * Replace mmd.boys (src/hf/core/core.f90:23-28, GSL 1F1) and rys.root5
 * (core.f90:232-812) in batched form.
 * sf_boys: F[i*(mmax+1)+m] = F_m(T[i]), 0 <= mmax <= 24.
 * sf_rys_roots: nodes as t^2 (the convention of root5's output,
 * core.f90:810-811), ascending, and weights; 1 <= nroots <= 7 (root5
 * has 5; 6 and 7 nodes are the exact rules for the f-shell quartets with
 * total angular momentum 10..13). */
int sf_boys(long long n, int mmax, const double *T, double *F, void *stream);
int sf_rys_roots(long long n, int nroots, const double *X, double *u, double *w,
                 void *stream);

/* ---- Contracted Cartesian-component quartets (host buffers) ------------
 * sf_eri_rys: one quartet with the argument meaning of rys.eri_rys
 * (core.f90:967-1002; called at src/hf/hf.py:26-30 and :100-104):
 * for function k (0..3), npgf[k] primitives stored back to back in coefs,
 * expns and norms (norms = the reference's normalize_factors), shells[3k..]
 * = (lx,ly,lz), centers[3k..] in bohr.  Total angular momentum L <= 13,
 * with floor(L/2) + 1 nodes: the reference's 5 nodes (core.f90:952) up to
 * L = 9, where they are exact; above, the exact 6- or 7-node rule (the
 * reference's 5-node sum is not exact there), status 5 beyond.
 * sf_eri_components: nq quartets over a table of nf such functions
 * (primitives of function f at [poff[f], poff[f]+npgf[f])); the batched
 * form of the Pool.map over _build_eri (hf.py:18-47). */
int sf_eri_rys(const int npgf[4], const double *coefs, const double *expns,
               const double *norms, const int shells[12], const double centers[12],
               double *out);
int sf_eri_components(int nf, const int *npgf, const int *poff, const double *coefs,
                      const double *expns, const double *norms, const int *shell,
                      const double *center, long long nq, const int *quartets,
                      double *out);

/* sf_eri_components without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks and
 * results as sf_eri_components (bit for bit); the function table, the
 * quartet list and the results are staged in a static 32 MiB device array
 * per GPU (held until the process ends), so one call holds at most ~1.3e6
 * quartets (each section 256-byte aligned: 4 nf + 4 nf + 24 nprim + 12 nf +
 * 24 nf + 24 nq bytes); larger calls return status 5.  Calls on one GPU
 * share the array: calls that overlap on one GPU race. */
int sf_eri_components_noalloc(int nf, const int *npgf, const int *poff, const double *coefs,
                              const double *expns, const double *norms, const int *shell,
                              const double *center, long long nq, const int *quartets,
                              double *out);

/* ---- Shell basis handle -------------------------------------------------
 * nshell contracted shells (L <= 3: s, p, d, f; status 5 above) in AO
 * order (f: xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz as
 * molecule.py:155-165).  Shell s: angular momentum
 * L[s], nprim[s] primitives (stored back to back in expns/coefs), centre
 * center[3s..3s+2] (bohr), atom shell_atom[s] in [0, natom) (status 1
 * otherwise; in a batch, one of its own molecule's nuclei).  shell_atom may
 * be NULL: then sf_grad refuses the handle (status 4).  coefs carry the contraction
 * coefficient times sqrt(2^(2L+1.5) a^(L+1.5) / pi^1.5), i.e. the
 * reference's normalize_factors*coefs of the x^L component times
 * sqrt((2L-1)!!); the library applies 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)
 * per component itself. */
int sf_basis_create(int nshell, const int *L, const int *nprim, const double *center,
                    const double *expns, const double *coefs, int natom,
                    const int *shell_atom, sf_basis **out);
int sf_basis_destroy(sf_basis *b);
int sf_basis_nbf(const sf_basis *b, int *nbf);

/* ---- Batches of independent molecules -----------------------------------
 * One handle for nmol molecules solved together (no reference counterpart:
 * the reference runs one molecule per scf_diis call, src/hf/hf.py:236-353).
 * Molecule m owns shells [mol_shell_begin[m], mol_shell_begin[m+1]) and
 * nuclei [mol_atom_begin[m], mol_atom_begin[m+1]) of the sf_basis_set_nuclei
 * list; shell arrays as in sf_basis_create.  No shell pair or quartet
 * crosses molecules; screening, pair significance and primitive-pair
 * screening are per molecule, so each molecule's integrals, J and K are
 * those of a single-molecule handle.  Every N x N matrix argument (P, J, K,
 * S, T, V, W) then holds one block per molecule: molecule m's n_m x n_m
 * block starts at element block_off[m] with leading dimension block_ld[m]
 * (default: blocks packed back to back, ld = n_m; set e.g. a padded
 * (nmol, nmax, nmax) layout with sf_basis_set_layout).  Elements outside
 * the blocks are never read or written.  For nmol = 1 this is exactly the
 * sf_basis_create handle.  sf_schwarz, sf_eri_dense and sf_grad need
 * nmol = 1. */
int sf_basis_create_batch(int nmol, const int *mol_shell_begin, int nshell, const int *L,
                          const int *nprim, const double *center, const double *expns,
                          const double *coefs, int natom, const int *mol_atom_begin,
                          const int *shell_atom, sf_basis **out);
/* Blocks must not overlap: [block_off[m], block_off[m] + n_m * block_ld[m]). */
int sf_basis_set_layout(sf_basis *b, const long long *block_off, const int *block_ld);
/* Current layout; *extent = elements a matrix buffer must span.  Output
 * arrays (nmol entries) may be NULL. */
int sf_basis_layout(const sf_basis *b, int *nmol, int *mol_nbf, long long *block_off,
                    int *block_ld, long long *extent);

/* Point charges for the nuclear-attraction integrals: Z[natom], xyz[3*natom]
 * (bohr).  Replaces the atom loop of hf.py:266-274 (mol.atoms[i].xyz, number). */
int sf_basis_set_nuclei(sf_basis *b, int natom, const double *Z, const double *xyz);
/* One-electron integrals S, T, V (device buffers in the handle's layout:
 * N x N row-major for one molecule; V with the -Z_C factor of the
 * molecule's own nuclei), drop-in for mmd2.overlap / kinetic / potential
 * summed as in hf.py:266-274 (mmd2.py:14-84). */
int sf_one_electron(sf_basis *b, double *S, double *T, double *V, void *stream);

/* Schwarz factors Q_ab = sqrt(max |(ij|ij)|), i in a, j in b, for every
 * shell pair a >= b, at Q[a*(a+1)/2 + b] (host buffer).  No counterpart in
 * the reference (which screens nothing). */
int sf_schwarz(sf_basis *b, double *Q);

/* Build the screened shell-quartet lists and the digestion plan: a quartet
 * survives when Q_ab * Q_cd >= tau (computed as that product in FP64);
 * in mixed precision, surviving quartets with Q_ab * Q_cd < tau_mixed run
 * in FP32.  tau = 0 keeps every quartet (the unscreened set of
 * build_eri_parallel, hf.py:37-43).  Replaces hf.py:33-61's enumeration. */
int sf_screen(sf_basis *b, double tau, double tau_mixed);
/* Primitive screening of the FP32 quartets (mixed precision only; no
 * reference counterpart): a primitive quartet whose primitive-pair Schwarz
 * product Qp_ab Qp_cd is below tau_prim is skipped.  Default (tau_prim < 0):
 * 1e-8 * tau_mixed, about the FP32 resolution of the largest FP32 quartet.
 * 0 disables it. */
int sf_set_prim_tau(sf_basis *b, double tau_prim);
/* Primitive screening of the FP64 quartets (no reference counterpart): a
 * primitive quartet with Qp_ab Qp_cd < tau_prim is skipped; each skipped
 * term is bounded by that product (Cauchy-Schwarz per primitive quartet),
 * so a quartet of np_ab x np_cd primitive pairs drops at most np_ab np_cd
 * tau_prim in magnitude.  Default (tau_prim < 0): 1e-3 * tau -- one skipped
 * term a thousandth of the shell-level Schwarz threshold (the sum stays
 * below tau for up to 1000 primitive quartets per quartet, e.g. every
 * quartet of 3-primitive s/p shells; contracted cc-pVDZ s shells can have
 * more); tau = 0 (no screening) therefore skips nothing.  0 disables it. */
int sf_set_prim_tau_fp64(sf_basis *b, double tau_prim);
/* counts[3*c + {0,1,2}] = surviving / FP64 / FP32 shell quartets of
 * quartet class c (0..54: classes 0..20 hold s, p, d shells, 21..54 an f
 * shell); *funcq = surviving function-level quartets. */
int sf_quartet_counts(const sf_basis *b, long long *counts, long long *funcq,
                      long long *primq);
/* Primitive quartets the Rys class kernels evaluate in one build of
 * `precision` under the current screening and primitive thresholds
 * (no reference counterpart; a measurement aid): counts[2c + seg], seg 0 =
 * the whole class (precision 0) or its FP64 segment, seg 1 = its FP32
 * segment (0 for precision 0).  Runs a counting kernel on the null stream. */
int sf_prim_quartet_counts(sf_basis *b, int precision, long long *counts);
/* The part of those primitive quartets that lies in far quartets (no
 * reference counterpart; a measurement aid): the (ss|ss) and (ps|ss)
 * quartets whose every primitive quartet has X = p q/(p+q) |P-Q|^2 >= 40,
 * which the library evaluates in the point-charge form S_b S_k / |P-Q|
 * (and its dipole term) instead of the Boys function; same layout as
 * sf_prim_quartet_counts, 0 for the other classes. */
int sf_far_prim_quartet_counts(sf_basis *b, int precision, long long *counts);
/* Sorted significant pairs of pair class pc (0..9: ss ps pp ds dp dd fs
 * fp fd ff): shells, Q and the
 * number of primitive pairs kept after primitive-pair screening (any output
 * pointer may be NULL). */
int sf_pair_list(const sf_basis *b, int pc, int *n, int *sa, int *sb, double *Q, int *nprim);
/* Per-bra ket counts of quartet class c (all, FP64 part). */
int sf_quartet_list(const sf_basis *b, int c, int *nbra, int *n_all, int *n_fp64);

/* Dense N^4 ERI tensor (host), drop-in for build_eri_parallel
 * (hf.py:33-61) after sf_screen(b, 0, 0). */
int sf_eri_dense(sf_basis *b, double *eri);

/* Direct-SCF J and K (device buffers in the handle's layout: N x N
 * row-major for one molecule, per-molecule blocks for a batch) for a symmetric
 * density P: J_pq = sum_rs (pq|rs) P_rs, K_pq = sum_rs (pr|qs) P_rs,
 * drop-in for the two einsums of hf.py:300 (G = J - K/2).
 * precision 0 = all FP64, 1 = mixed (FP32 below tau_mixed, FP64
 * accumulation).  Deterministic: no atomics. */
int sf_jk(sf_basis *b, const double *P, double *J, double *K, int precision, void *stream);
/* sf_jk with HIP events between launch groups on `stream` (for measurement;
 * no counterpart in the reference, whose only timer is utils.timeit,
 * src/hf/utils.py:76-83).  ms[112] receives milliseconds: ms[2c] / ms[2c+1]
 * = FP64 / FP32 ERI+digestion kernel of quartet class c (for a class with
 * more than 54 integrals per quartet in a mixed build, whose two segments
 * share one digestion launch: its integrals / the FP32 segment's integrals
 * plus that digestion; c = 0..54), ms[110] = density scaling, ms[111] =
 * deterministic plan reduction and finalisation. */
int sf_jk_profile(sf_basis *b, const double *P, double *J, double *K, int precision,
                  void *stream, double *ms);
/* sf_jk through a captured hipGraph: the first call (and the first after
 * the arguments, sf_screen, sf_autotune, sf_set_tuning or
 * sf_basis_set_layout change) captures the whole build -- density scaling,
 * every quartet-class kernel, the plan reduction and finalisation, kernel
 * nodes only -- on a private stream; every call replays it on `stream`.
 * Same results as sf_jk, bit for bit.  For SCF loops that rebuild J/K for
 * fixed buffers (the per-iteration line hf.py:300). */
int sf_jk_graph(sf_basis *b, const double *P, double *J, double *K, int precision,
                void *stream);
/* Launch-shape autotuner (no reference counterpart).  Three tuning slots
 * per quartet class c: 0 = the whole class in FP64 builds (precision 0),
 * 1 / 2 = the FP64 / FP32 segment of mixed builds (precision 1).  For the
 * slots of `precision`, times 64/128/256-thread blocks with one thread per
 * quartet or persistent grids of 2 or 8 blocks per compute unit on the
 * current screened lists (density P, device; each candidate timed on the
 * first 2^20 quartets of the segment) and keeps the fastest; for slot 2 it
 * also times the segment with FP64 arithmetic (same primitive threshold;
 * never less accurate) and keeps that where it is faster; a class whose
 * FP32 segment runs in FP64 arithmetic then builds both segments in one
 * launch (slot 1's shape), each quartet with its own segment's primitive
 * threshold.  sf_jk and
 * sf_jk_graph use the choice.  shapes[9c + 3slot + {0,1,2}] receives
 * (block, grid; 0 = one per quartet, arith: 1 = FP64 arithmetic in slot 2)
 * for every slot, ms[3c + slot] the chosen time of each tuned slot (others
 * untouched); either may be NULL. */
int sf_autotune(sf_basis *b, const double *P, int precision, void *stream, int *shapes,
                double *ms);
/* Set the shapes (all 3 slots of all 55 classes, 495 ints) directly, in
 * the layout sf_autotune reports. */
int sf_set_tuning(sf_basis *b, const int *shapes);
/* Integral engine of quartet class qcls (0..54, or -1 for all): 0 = Rys
 * quadrature (default; the method of rys.eri_rys, core.f90:917-1002),
 * 1 = McMurchie-Davidson (the method of mmd.eri, core.f90:100-198) with the
 * Hermite-to-Cartesian E.R.E transforms on the FP64 matrix cores.  Engine 1
 * computes in FP64 in both precision segments and covers the s/p/d classes
 * 0..20 (status 5 for an f-shell class; qcls = -1 leaves those on Rys).
 * Used by sf_jk, sf_jk_graph, sf_autotune and sf_eri_dense. */
int sf_set_engine(sf_basis *b, int qcls, int engine);
/* Side streams of sf_jk / sf_jk_graph (0..3; default 2 for sf_jk and 3 for
 * the build sf_jk_graph captures, the fastest of each measured): the quartet-class
 * launches, and then the per-tile-shape reductions, are spread over the
 * caller's stream and nside private streams (forked from and joined back
 * into the caller's stream with events), so small classes run beside large
 * ones.  0 = every launch on the caller's stream, in class order.  Results
 * are the same either way (every launch writes its own slots).  Launch
 * scheduling only; no counterpart in the reference. */
int sf_set_streams(sf_basis *b, int nside);

/* Nuclear gradient of the RHF electronic energy (host output grad[3*natom],
 * hartree/bohr, atom order of the basis): sum P dH/dR - sum W dS/dR +
 * 1/2 sum Gamma d(pq|rs)/dR for the density P and energy-weighted density
 * W = 2 sum_occ e_i C_i C_i^T (device buffers, N x N).  Derivative ERIs
 * over the screened quartet list of sf_screen, deterministic reductions.
 * precision 0 = all FP64; 1 = derivative quartets below tau_mixed in FP32
 * (nodes, weights and every sum in FP64), as sf_jk, or -- per class, where
 * the first mixed call after sf_screen measures it faster -- in FP64
 * arithmetic within one launch over the class.  Primitive quartets below
 * the segment's threshold are skipped as in sf_jk (sf_set_prim_tau_fp64,
 * sf_set_prim_tau).
 * No counterpart in the reference (single-point energies only); nuclear
 * repulsion is not included.  Requires sf_basis_set_nuclei for every atom.
 * s, p, d shells: a basis with f shells returns status 5. */
int sf_grad(sf_basis *b, const double *P, const double *W, double *grad, int precision,
            void *stream);

#ifdef __cplusplus
}
#endif
#endif
