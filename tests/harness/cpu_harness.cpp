// CPU harness: compiles the HIP kernels' host+device math headers with g++
// so that tests can check them against the oracle without a GPU.
// Test infrastructure only.
#include "../../segfault_py_amd/csrc/eri_gen.hpp"
#include "../../segfault_py_amd/csrc/mmd.hpp"

#include <vector>

using namespace sf;

extern "C" void h_boys(int mmax, double X, double* F) { boys_grid(mmax, X, F); }
extern "C" void h_boys01(double X, double* F) { boys01(X, F[0], F[1]); }
// F_0, F_1 as the one-node kernels evaluate them: from the LDS Taylor rows
// (built here from the same element function the kernels stage).
extern "C" void h_boys01_kernel(double X, double* F)
{
    static double tab[node_table_lds(1)];
    static bool init = false;
    if (!init) {
        for (int k = 0; k < node_table_lds(1); ++k) tab[k] = node_table_elem<1>(k);
        init = true;
    }
    boys01_node(X, NodeTable{tab, kBoysTaylorCols, 1}, F[0], F[1]);
}
// The node evaluation the ERI kernels use (one node from F_0, F_1 of the
// Boys grid, more from the Chebyshev tables), with the global tables.
extern "C" int h_rys_kernel(int n, double X, double* u, double* w)
{
    switch (n) {
    case 1: rys_nodes_tab<1>(X, global_node_table<1>(), u, w); return 0;
    case 2: rys_nodes_tab<2>(X, global_node_table<2>(), u, w); return 0;
    case 3: rys_nodes_tab<3>(X, global_node_table<3>(), u, w); return 0;
    case 4: rys_nodes_tab<4>(X, global_node_table<4>(), u, w); return 0;
    case 5: rys_nodes_tab<5>(X, global_node_table<5>(), u, w); return 0;
    case 6: rys_nodes_tab<6>(X, global_node_table<6>(), u, w); return 0;
    case 7: rys_nodes_tab<7>(X, global_node_table<7>(), u, w); return 0;
    }
    return 1;
}

extern "C" int h_rys(int n, double X, double* u, double* w)
{
    switch (n) {
    case 1: rys_roots<1>(X, u, w); return 0;
    case 2: rys_roots<2>(X, u, w); return 0;
    case 3: rys_roots<3>(X, u, w); return 0;
    case 4: rys_roots<4>(X, u, w); return 0;
    case 5: rys_roots<5>(X, u, w); return 0;
    case 6: rys_roots<6>(X, u, w); return 0;
    case 7: rys_roots<7>(X, u, w); return 0;
    }
    return 1;
}

template <int LA, int LB, int LC, int LD>
static void run(const double* bra, int nb, const double* ket, int nk, const double* A,
                const double* B, const double* C, const double* D, double* out, int fp32)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    if (fp32) {
        float acc[S::NINT] = {};
        eri_shell_quartet<LA, LB, LC, LD, float>((const PrimPair*)bra, nb, (const PrimPair*)ket, nk,
                                                 A, B, C, D, acc);
        for (int i = 0; i < S::NINT; ++i) out[i] = acc[i];
    } else {
        double acc[S::NINT] = {};
        eri_shell_quartet<LA, LB, LC, LD, double>((const PrimPair*)bra, nb, (const PrimPair*)ket, nk,
                                                  A, B, C, D, acc);
        for (int i = 0; i < S::NINT; ++i) out[i] = acc[i];
    }
}

// The split kernels' decomposition: per component of A (eri_bra_row) where
// the class allows it, else per bra component pair; pairs = 1 forces the
// per-pair path.
template <int LA, int LB, int LC, int LD>
static void run_split(const double* bra, int nb, const double* ket, int nk, const double* A,
                      const double* B, const double* C, const double* D, double* out, int fp32,
                      int pairs = 0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    if constexpr (S::SPLIT_ROWS) {
        if (!pairs) {
            constexpr int NO = S::NB * S::NC * S::ND;
            for (int s = 0; s < S::NA; ++s) {
                static_switch<0, S::NA>(s, [&](auto sc) {
                    constexpr int FA = decltype(sc)::value;
                    if (fp32) {
                        float acc[NO] = {};
                        eri_bra_row<LA, LB, LC, LD, FA, float>((const PrimPair*)bra, nb,
                                                               (const PrimPair*)ket, nk, A, B, C, D, acc);
                        for (int k = 0; k < NO; ++k) out[s * NO + k] = acc[k];
                    } else {
                        double acc[NO] = {};
                        eri_bra_row<LA, LB, LC, LD, FA, double>((const PrimPair*)bra, nb,
                                                                (const PrimPair*)ket, nk, A, B, C, D, acc);
                        for (int k = 0; k < NO; ++k) out[s * NO + k] = acc[k];
                    }
                });
            }
            return;
        }
    }
    for (int s = 0; s < S::NA * S::NB; ++s) {
        static_switch<0, S::NA * S::NB>(s, [&](auto sc) {
            constexpr int FA = decltype(sc)::value / S::NB, FB = decltype(sc)::value % S::NB;
            if (fp32) {
                float acc[S::NC * S::ND] = {};
                eri_bra_component<LA, LB, LC, LD, FA, FB, float>(
                    (const PrimPair*)bra, nb, (const PrimPair*)ket, nk, A, B, C, D, acc);
                for (int k = 0; k < S::NC * S::ND; ++k) out[s * S::NC * S::ND + k] = acc[k];
            } else {
                double acc[S::NC * S::ND] = {};
                eri_bra_component<LA, LB, LC, LD, FA, FB, double>(
                    (const PrimPair*)bra, nb, (const PrimPair*)ket, nk, A, B, C, D, acc);
                for (int k = 0; k < S::NC * S::ND; ++k) out[s * S::NC * S::ND + k] = acc[k];
            }
        });
    }
}

static QuartetMap dense_map(int nbf, const int* f0)
{
    QuartetMap qm;
    for (int k = 0; k < 4; ++k) {
        qm.row[k] = (long long)f0[k] * nbf;
        qm.col[k] = f0[k];
    }
    qm.ld = nbf;
    return qm;
}

template <int LA, int LB, int LC, int LD>
static void run_tiles(const double* v, double fac, const double* P, int nbf, const int* f0,
                      double* t)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    const QuartetMap qm = dense_map(nbf, f0);
    digest_tile<LA, LB, LC, LD, 0>(v, 1, fac, P, qm, t + S::T_AB);
    digest_tile<LA, LB, LC, LD, 1>(v, 1, fac, P, qm, t + S::T_CD);
    digest_tile<LA, LB, LC, LD, 2>(v, 1, fac, P, qm, t + S::T_AC);
    digest_tile<LA, LB, LC, LD, 3>(v, 1, fac, P, qm, t + S::T_AD);
    digest_tile<LA, LB, LC, LD, 4>(v, 1, fac, P, qm, t + S::T_BC);
    digest_tile<LA, LB, LC, LD, 5>(v, 1, fac, P, qm, t + S::T_BD);
}

// The same six tiles digested group by group (digest_tile_group over the
// TileGroups cuts, the split-class digestion's tile mode).
template <int LA, int LB, int LC, int LD>
static void run_tile_groups(const double* v, double fac, const double* P, int nbf, const int* f0,
                            double* t)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    using TG = TileGroups<LA, LB, LC, LD>;
    const QuartetMap qm = dense_map(nbf, f0);
    double* const dst[6] = {t + S::T_AB, t + S::T_CD, t + S::T_AC,
                            t + S::T_AD, t + S::T_BC, t + S::T_BD};
    static_for<0, TG::count>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
        digest_tile_group<LA, LB, LC, LD, TG::begin(g), TG::end(g)>(v, 1, fac, P, qm, dst);
    });
}

// The same six tiles built row by row (digest_tile_row, the split-class
// digestion's row mode).
template <int LA, int LB, int LC, int LD, int T>
static void tile_by_rows(const double* v, double fac, const double* P, const QuartetMap& qm,
                         double* t)
{
    using G = TileGeom<LA, LB, LC, LD, T>;
    double Pm[G::M1][G::M2];
    load_tile_density<LA, LB, LC, LD, T>(P, qm, Pm);
    for (int r = 0; r < G::N1; ++r) digest_tile_row<LA, LB, LC, LD, T>(v, 1, fac, Pm, r, t + r * G::N2);
}

template <int LA, int LB, int LC, int LD>
static void run_tile_rows(const double* v, double fac, const double* P, int nbf, const int* f0,
                          double* t)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    const QuartetMap qm = dense_map(nbf, f0);
    tile_by_rows<LA, LB, LC, LD, 0>(v, fac, P, qm, t + S::T_AB);
    tile_by_rows<LA, LB, LC, LD, 1>(v, fac, P, qm, t + S::T_CD);
    tile_by_rows<LA, LB, LC, LD, 2>(v, fac, P, qm, t + S::T_AC);
    tile_by_rows<LA, LB, LC, LD, 3>(v, fac, P, qm, t + S::T_AD);
    tile_by_rows<LA, LB, LC, LD, 4>(v, fac, P, qm, t + S::T_BC);
    tile_by_rows<LA, LB, LC, LD, 5>(v, fac, P, qm, t + S::T_BD);
}

template <int LA, int LB, int LC, int LD>
static void run_digest(const double* v, const double* P, int nbf, const int* f0, double* t)
{
    digest_quartet<LA, LB, LC, LD, double>(v, P, dense_map(nbf, f0), t);
}

#define SF_CLASSES(X)                                                                        \
    X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1)       \
    X(2, 0, 0, 0) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0)     \
    X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1)     \
    X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)

extern "C" int h_eri(int la, int lb, int lc, int ld, const double* bra, int nb, const double* ket, int nk,
          const double* A, const double* B, const double* C, const double* D, double* out, int fp32)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, out, fp32);              \
        return 0;                                                              \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_digest(int la, int lb, int lc, int ld, const double* v, const double* P, int nbf,
             const int* f0, double* t)
{
#define X(a, b, c, d)                                      \
    if (la == a && lb == b && lc == c && ld == d) {        \
        run_digest<a, b, c, d>(v, P, nbf, f0, t);          \
        return 0;                                          \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_eri_split(int la, int lb, int lc, int ld, const double* bra, int nb,
                           const double* ket, int nk, const double* A, const double* B,
                           const double* C, const double* D, double* out, int fp32)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run_split<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, out, fp32);        \
        return 0;                                                              \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_eri_split_pairs(int la, int lb, int lc, int ld, const double* bra, int nb,
                                 const double* ket, int nk, const double* A, const double* B,
                                 const double* C, const double* D, double* out, int fp32)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run_split<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, out, fp32, 1);     \
        return 0;                                                              \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_digest_tiles(int la, int lb, int lc, int ld, const double* v, double fac,
                              const double* P, int nbf, const int* f0, double* t)
{
#define X(a, b, c, d)                                      \
    if (la == a && lb == b && lc == c && ld == d) {        \
        run_tiles<a, b, c, d>(v, fac, P, nbf, f0, t);      \
        return 0;                                          \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_digest_tile_groups(int la, int lb, int lc, int ld, const double* v, double fac,
                                    const double* P, int nbf, const int* f0, double* t)
{
#define X(a, b, c, d)                                        \
    if (la == a && lb == b && lc == c && ld == d) {          \
        run_tile_groups<a, b, c, d>(v, fac, P, nbf, f0, t);  \
        return 0;                                            \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_digest_tile_rows(int la, int lb, int lc, int ld, const double* v, double fac,
                                  const double* P, int nbf, const int* f0, double* t)
{
#define X(a, b, c, d)                                      \
    if (la == a && lb == b && lc == c && ld == d) {        \
        run_tile_rows<a, b, c, d>(v, fac, P, nbf, f0, t);  \
        return 0;                                          \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_one_electron(int la, int lb, const double* prim, int np, const double* A,
                              const double* B, int natom, const double* Z, const double* C,
                              double* s, double* t, double* v)
{
    const PrimPair* pp = (const PrimPair*)prim;
#define O_(x, y)                                                                     \
    if (la == x && lb == y) {                                                        \
        one_electron_shell_pair<x, y>(pp, np, A, B, natom, Z, C, s, t, v);           \
        return 0;                                                                    \
    }
    O_(0, 0) O_(1, 0) O_(1, 1) O_(2, 0) O_(2, 1) O_(2, 2) O_(3, 0) O_(3, 1) O_(3, 2) O_(3, 3)
#undef O_
    return 1;
}

// The run-time-L path of the f-shell classes (k_eri_gen.hip): every
// component of one shell quartet, component by component (gen_component,
// no primitive screening), in (a, b, c, d) order.
extern "C" int h_eri_gen(int la, int lb, int lc, int ld, const double* bra, int nb,
                         const double* ket, int nk, const double* A, const double* B,
                         const double* C, const double* D, double* out)
{
    const GenClass g = gen_class(la, lb, lc, ld);
    if (g.nr > SF_RYS_NMAX) return 1;
    for (int k = 0; k < g.nint; ++k) {
        int lcx[12];
        gen_exps(g, k, lcx);
        double v = 0.0;
        static_switch<1, SF_RYS_NMAX + 1>(g.nr, [&](auto nc) {
            constexpr int NR = decltype(nc)::value;
            v = gen_component<NR>(lcx, A, B, C, D, (const PrimPair*)bra, nb, (const PrimPair*)ket,
                                  nk, nullptr, nullptr, 0.0);
        });
        out[k] = v;
    }
    return 0;
}

// The large f classes' path (k_eri_gen<LA, LB, LC, LD>): every bra
// component pair at run time through eri_bra_runtime, no primitive
// screening, integrals in (a, b, c, d) order.
template <int LA, int LB, int LC, int LD>
static void run_bra_runtime(const double* bra, int nb, const double* ket, int nk, const double* A,
                            const double* B, const double* C, const double* D, double* out)
{
    constexpr int NB = cart_n(LB), NK = cart_n(LC) * cart_n(LD);
    for (int s = 0; s < cart_n(LA) * NB; ++s) {
        double acc[NK] = {};
        eri_bra_runtime<LA, LB, LC, LD>(s / NB, s % NB, (const PrimPair*)bra, nb,
                                        (const PrimPair*)ket, nk, A, B, C, D, acc, nullptr,
                                        nullptr, 0.0);
        for (int k = 0; k < NK; ++k) out[s * NK + k] = acc[k];
    }
}

extern "C" int h_eri_bra_runtime(int la, int lb, int lc, int ld, const double* bra, int nb,
                                 const double* ket, int nk, const double* A, const double* B,
                                 const double* C, const double* D, double* out)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run_bra_runtime<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, out);        \
        return 0;                                                              \
    }
    X(3, 1, 2, 2) X(3, 1, 3, 1) X(3, 2, 2, 1) X(3, 2, 2, 2) X(3, 2, 3, 0) X(3, 2, 3, 1)
    X(3, 2, 3, 2) X(3, 3, 1, 1) X(3, 3, 2, 0) X(3, 3, 2, 1) X(3, 3, 2, 2) X(3, 3, 3, 0)
    X(3, 3, 3, 1) X(3, 3, 3, 2) X(3, 3, 3, 3) X(2, 1, 1, 0) X(1, 0, 0, 0)
#undef X
    return 1;
}

// The six digestion tiles of one quartet element by element
// (gen_tile_element, as k_digest_gen), laid out as digest_quartet's t[].
extern "C" int h_gen_tiles(int la, int lb, int lc, int ld, const double* v, double fac,
                           const double* P, int nbf, const int* f0, double* t)
{
    const GenClass g = gen_class(la, lb, lc, ld);
    const QuartetMap qm = dense_map(nbf, f0);
    for (int T = 0; T < 6; ++T) {
        const int n = (T == 5 ? g.ntile : g.toff[T + 1]) - g.toff[T];
        for (int e = 0; e < n; ++e) t[g.toff[T] + e] = gen_tile_element(g, T, e, v, 1, fac, P, qm);
    }
    return 0;
}

// The kernels' decomposition: every component of A at once (grad_all_a)
// where the class allows it, else per bra component pair; slices = 1
// forces the per-pair path.
template <int LA, int LB, int LC, int LD>
static void run_grad(const double* bra, int nb, const double* ket, int nk, const double* A,
                     const double* B, const double* C, const double* D, const double* Dfac,
                     double* g, int slices = 0)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    for (int k = 0; k < 9; ++k) g[k] = 0.0;
    if constexpr (S::GRAD_ALL_A) {
        if (!slices) {
            grad_all_a<LA, LB, LC, LD, double>((const PrimPair*)bra, nb, (const PrimPair*)ket, nk, A, B,
                                               C, D, Dfac, g,
                                               global_node_table<(S::L + 1) / 2 + 1>());
            return;
        }
    }
    for (int s = 0; s < S::NA * S::NB; ++s) {
        static_switch<0, S::NA * S::NB>(s, [&](auto sc) {
            constexpr int FA = decltype(sc)::value / S::NB, FB = decltype(sc)::value % S::NB;
            grad_bra_component<LA, LB, LC, LD, FA, FB, double>(
                (const PrimPair*)bra, nb, (const PrimPair*)ket, nk, A, B, C, D,
                Dfac + s * S::NC * S::ND, g);
        });
    }
}

extern "C" int h_grad(int la, int lb, int lc, int ld, const double* bra, int nb, const double* ket,
                      int nk, const double* A, const double* B, const double* C, const double* D,
                      const double* Dfac, double* g)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run_grad<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, Dfac, g);           \
        return 0;                                                              \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_grad_slices(int la, int lb, int lc, int ld, const double* bra, int nb,
                             const double* ket, int nk, const double* A, const double* B,
                             const double* C, const double* D, const double* Dfac, double* g)
{
#define X(a, b, c, d)                                                          \
    if (la == a && lb == b && lc == c && ld == d) {                            \
        run_grad<a, b, c, d>(bra, nb, ket, nk, A, B, C, D, Dfac, g, 1);        \
        return 0;                                                              \
    }
    SF_CLASSES(X)
#undef X
    return 1;
}

extern "C" int h_grad_onee(int la, int lb, const double* prim, int np, const double* A,
                           const double* B, int atomA, int atomB, int natom, const double* Z,
                           const double* C, const double* Pb, const double* Wb, double mult,
                           double* row)
{
    const PrimPair* pp = (const PrimPair*)prim;
#define O_(x, y)                                                                              \
    if (la == x && lb == y) {                                                                 \
        one_electron_grad_pair<x, y>(pp, np, A, B, atomA, atomB, natom, Z, C, Pb, Wb, mult, row); \
        return 0;                                                                             \
    }
    O_(0, 0) O_(1, 0) O_(1, 1) O_(2, 0) O_(2, 1) O_(2, 2)
#undef O_
    return 1;
}

// The one-node classes' point-charge form for far quartets
// (eri_one_node_far, no primitive screening); bs / ks: S of each primitive
// pair as the library computes it, pi^{3/2} (K/p) sqrt(1/p).
extern "C" int h_eri_far(int la, const double* bra, int nb, const double* ket, int nk,
                         const double* A, double* out)
{
    std::vector<double> bs(nb), ks(nk);
    const PrimPair* pb = (const PrimPair*)bra;
    const PrimPair* pk = (const PrimPair*)ket;
    for (int k = 0; k < nb; ++k) bs[k] = 5.568327996831708 * pb[k].Kp * std::sqrt(pb[k].ip);
    for (int k = 0; k < nk; ++k) ks[k] = 5.568327996831708 * pk[k].Kp * std::sqrt(pk[k].ip);
    double acc[3] = {0.0, 0.0, 0.0};
    if (la == 0)
        eri_one_node_far<0, double, false, const PrimPair*, const double*, const double*>(
            pb, bs.data(), nb, pk, ks.data(), nk, A, acc, nullptr, nullptr, 0.0);
    else if (la == 1)
        eri_one_node_far<1, double, false, const PrimPair*, const double*, const double*>(
            pb, bs.data(), nb, pk, ks.data(), nk, A, acc, nullptr, nullptr, 0.0);
    else
        return 1;
    for (int k = 0; k < (la == 0 ? 1 : 3); ++k) out[k] = acc[k];
    return 0;
}
