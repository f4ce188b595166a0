// k_eri.hip -- per-angular-momentum-class ERI kernels for gfx950.
//
// One thread computes one contracted shell quartet (ab|cd) of its class
// (all primitive quartets, all Cartesian components, nodes floor(L/2)+1)
// and either
//   * digests it into six J/K partial tiles (kModeJK), written as SoA planes
//     out[tile_element * qstride + q] so a wavefront's stores are contiguous,
//   * or stores the raw component integrals (kModeRaw, dense export).
// The flat quartet index of a segment is mapped back to (bra, ket) by a
// binary search of the per-bra prefix sums; consecutive threads share the
// bra pair, so its primitive data is read once per wavefront from cache.
//
// Replaces the per-quartet Python fan-out + f2py call of hf.py:18-61 and
// the einsum digestion of hf.py:300; the integral math follows
// rys.eri_rys (src/hf/core/core.f90:917-1002) -- see eri_core.hpp.
#include "eri_mmd.hpp"
#include "sf_internal.hpp"

// Built once per quartet class with -DSF_CLASS=<id> (that class's kernels
// only, so the classes compile in parallel) and once without SF_CLASS (the
// dispatch tables).
#define SF_FOR_EACH_CLASS(X)                                                                     \
    X(0, 0, 0, 0, 0) X(1, 1, 0, 0, 0) X(2, 1, 0, 1, 0) X(3, 1, 1, 0, 0) X(4, 1, 1, 1, 0)          \
    X(5, 1, 1, 1, 1) X(6, 2, 0, 0, 0) X(7, 2, 0, 1, 0) X(8, 2, 0, 1, 1) X(9, 2, 0, 2, 0)          \
    X(10, 2, 1, 0, 0) X(11, 2, 1, 1, 0) X(12, 2, 1, 1, 1) X(13, 2, 1, 2, 0) X(14, 2, 1, 2, 1)    \
    X(15, 2, 2, 0, 0) X(16, 2, 2, 1, 0) X(17, 2, 2, 1, 1) X(18, 2, 2, 2, 0) X(19, 2, 2, 2, 1)    \
    X(20, 2, 2, 2, 2)

// f-shell classes with compiled kernels (J/K and the dense export; the
// gradient and the McMurchie-Davidson engine cover s, p, d): those with at
// most 54 integrals per quartet (one thread per quartet) or at most 540 and
// at most 54 per component of A (one wavefront per (64 quartets, component
// of A)).  Each is built in a unit of its own (-DSF_FCLASS=id with
// -DSF_FLA..SF_FLD); compile time grows steeply beyond (measured: (fd|pp)
// 43 s, (fd|dp) 6 min, (ff|ff) 26 min), so the 15 larger f classes run the
// run-time-L kernels of k_eri_gen.hip.
#define SF_FOR_EACH_FCLASS(X)                                                                   \
    X(21, 3, 0, 0, 0) X(22, 3, 0, 1, 0) X(23, 3, 0, 1, 1) X(24, 3, 0, 2, 0) X(25, 3, 0, 2, 1)   \
    X(26, 3, 0, 2, 2) X(27, 3, 0, 3, 0) X(28, 3, 1, 0, 0) X(29, 3, 1, 1, 0) X(30, 3, 1, 1, 1)   \
    X(31, 3, 1, 2, 0) X(32, 3, 1, 2, 1) X(34, 3, 1, 3, 0) X(36, 3, 2, 0, 0) X(37, 3, 2, 1, 0)   \
    X(38, 3, 2, 1, 1) X(39, 3, 2, 2, 0) X(45, 3, 3, 0, 0) X(46, 3, 3, 1, 0)

namespace sf {

using SchwarzLaunch = int (*)(const int*, const int*, const int*, const int*, int, const PrimPair*,
                              const double*, double*, double*, hipStream_t);
#define X(id, a, b, c, d)                                                                       \
    int launch_class_##id(int fp32, int mode, const ClassParams& cp, const Tune& tn,           \
                          hipStream_t s);                                                      \
    int launch_schwarz_##id(const int*, const int*, const int*, const int*, int,               \
                            const PrimPair*, const double*, double*, double*, hipStream_t);    \
    int launch_grad_##id(int fp32, const ClassParams& cp, double* gout, hipStream_t s);        \
    int launch_mmd_##id(int mode, const ClassParams& cp, const Tune& tn, hipStream_t s);
SF_FOR_EACH_CLASS(X)
#undef X
#define X(id, a, b, c, d)                                                                       \
    int launch_class_##id(int fp32, int mode, const ClassParams& cp, const Tune& tn,           \
                          hipStream_t s);
SF_FOR_EACH_FCLASS(X)
#undef X

#ifdef SF_CLASS

#define SF_TRY_K(x)                   \
    do {                              \
        const int rc_ = (x);          \
        if (rc_ != kOk) return rc_;   \
    } while (0)

// Copies the node table of NR nodes into LDS (block-wide; every thread of
// the block must call it) and returns its descriptor; global table when
// the class has no LDS copy.
template <int NR>
__device__ __forceinline__ NodeTable stage_node_table(double* s_tab)
{
    constexpr int n = node_table_lds(NR);
    if constexpr (n > 0) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) s_tab[k] = node_table_elem<NR>(k);
        __syncthreads();
        return NodeTable{s_tab, NR == 1 ? kBoysTaylorCols : 0, NR == 1 ? 1 : 0};
    } else {
        return global_node_table<NR>();
    }
}

__device__ __forceinline__ QuartetMap quartet_map(const ShellMap& m, int a, int b, int c, int d)
{
    QuartetMap qm;
    qm.row[0] = m.row[a];
    qm.row[1] = m.row[b];
    qm.row[2] = m.row[c];
    qm.row[3] = m.row[d];
    qm.col[0] = m.col[a];
    qm.col[1] = m.col[b];
    qm.col[2] = m.col[c];
    qm.col[3] = m.col[d];
    qm.ld = m.ld[a];
    return qm;
}

// One shell quartet (bra i, ket j, class-wide index qa): integrals, then
// digestion tiles or raw planes.  When the caller passes a wave-uniform i
// (readfirstlane), the bra pair's primitives, centres and density rows are
// fetched with scalar loads.
template <int LA, int LB, int LC, int LD, typename R, int MODE, bool UNI>
__device__ __forceinline__ void class_quartet(const ClassParams& cp, const NodeTable& nt, int i,
                                              int j, long long qa)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    if constexpr (UNI) i = __builtin_amdgcn_readfirstlane(i);
    const int a = cp.bra.sa[i], b = cp.bra.sb[i];
    const int c = cp.ket.sa[j], d = cp.ket.sb[j];
    const double A[3] = {cp.shell_xyz[3 * a], cp.shell_xyz[3 * a + 1], cp.shell_xyz[3 * a + 2]};
    const double B[3] = {cp.shell_xyz[3 * b], cp.shell_xyz[3 * b + 1], cp.shell_xyz[3 * b + 2]};
    // the ket's centres and primitive pairs from the transposed arrays:
    // consecutive lanes (consecutive kets j) read consecutive doubles
    const long long kn = cp.kt_n;
    const double C[3] = {cp.kc[j], cp.kc[kn + j], cp.kc[2 * kn + j]};
    const double D[3] = {cp.kc[3 * kn + j], cp.kc[4 * kn + j], cp.kc[5 * kn + j]};
    const KetView kv{cp.kt + j, cp.kt_fs, kn};
    const StridedView kqv{cp.kt + kKetQ * cp.kt_fs + j, kn};
    R acc[S::NINT];
#pragma unroll
    for (int k = 0; k < S::NINT; ++k) acc[k] = (R)0;
    // primitive quartets with Qp_b Qp_k < tau are skipped (the segment's
    // threshold; 0 keeps every one); in a launch over both mixed segments,
    // the bra's kets past its n64 FP64 ones take the FP32 segment's threshold
    constexpr bool PS = true;
    double tau = cp.tau_prim;
    if (cp.n64 && j - cp.ket0[i] >= cp.n64[i]) tau = cp.tau_prim_lo;
    int bpo = cp.bra.poff[i], bnp = cp.bra.np[i];
    // one-node classes, far quartets: the point-charge form
    constexpr bool far = MODE == kModeJKFar;
    static_assert(!far || S::L <= 1, "the point-charge form is for one-node classes");
    const int knp = cp.ket.np[j];
    if constexpr (UNI) {
        // the bra's primitive pairs are read-only for the whole launch: read
        // them through the constant address space, i.e. with scalar loads
        bpo = __builtin_amdgcn_readfirstlane(bpo);
        bnp = __builtin_amdgcn_readfirstlane(bnp);
        using CPrim = const __attribute__((address_space(4))) PrimPair*;
        using CDbl = const __attribute__((address_space(4))) double*;
        if constexpr (far) {
            if constexpr (S::L <= 1)
                eri_one_node_far<LA, R, PS, CPrim, CDbl, CDbl>(
                    (CPrim)(cp.prim + bpo), (CDbl)(cp.prims + bpo), bnp, kv,
                    StridedView{cp.kt + kKetS * cp.kt_fs + j, kn}, knp, A, acc,
                    (CDbl)(cp.primq + bpo), kqv, tau);
        } else {
            eri_shell_quartet<LA, LB, LC, LD, R, PS, CPrim, CDbl>(
                (CPrim)(cp.prim + bpo), bnp, kv, knp, A, B, C, D, acc, nt,
                (CDbl)(cp.primq + bpo), kqv, tau);
        }
    } else {
        if constexpr (far) {
            if constexpr (S::L <= 1)
                eri_one_node_far<LA, R, PS, const PrimPair*, const double*, const double*>(
                    cp.prim + bpo, cp.prims + bpo, bnp, kv,
                    StridedView{cp.kt + kKetS * cp.kt_fs + j, kn}, knp, A, acc, cp.primq + bpo,
                    kqv, tau);
        } else {
            eri_shell_quartet<LA, LB, LC, LD, R, PS>(cp.prim + bpo, bnp, kv, knp, A, B, C, D, acc,
                                                     nt, cp.primq + bpo, kqv, tau);
        }
    }
    if constexpr (MODE == kModeRaw) {
#pragma unroll
        for (int k = 0; k < S::NINT; ++k) cp.out[k * cp.qstride + qa] = (double)acc[k];
    } else {
        const double fac = ((a == b) ? 0.5 : 1.0) * ((c == d) ? 0.5 : 1.0) *
                           ((cp.same && i == j) ? 0.5 : 1.0);
        double v[S::NINT];
#pragma unroll
        for (int k = 0; k < S::NINT; ++k) v[k] = fac * (double)acc[k];
        double t[S::NTILE];
        digest_quartet<LA, LB, LC, LD, double>(v, cp.P, quartet_map(cp.map, a, b, c, d), t);
        // each tile goes to its destination-sorted slot (contiguous elements)
#pragma unroll
        for (int T = 0; T < 6; ++T) {
            double* dst = cp.tiles + cp.pos[T * cp.qstride + qa];
#pragma unroll
            for (int e = 0; e < S::tile_n1(T) * S::tile_n2(T); ++e) dst[e] = t[S::tile_off(T) + e];
        }
    }
}

// Minimum wavefronts per SIMD the register allocator must leave room for,
// per class (1 = no constraint).  As one thread per quartet (ds|pp) and
// (dp|ps) (54 integrals per quartet) need more than 256 registers per lane,
// i.e. one wavefront per SIMD; held to two they spill ~64 registers to
// scratch and still run faster (20-water build: 0.305 -> 0.268 ms and
// 0.226 -> 0.189 ms), as (ds|ps) held to three (175 -> 168 registers,
// 0.405 -> 0.380 ms).  The same for (ps|ps), (pp|ss) (held to four) and
// (pp|ps) (three) measured slower; they keep the allocator's choice.
template <int LA, int LB, int LC, int LD>
struct ClassWaves {
    static constexpr int code = LA * 1000 + LB * 100 + LC * 10 + LD;
    static constexpr int min = (code == 2110 || code == 2011) ? 2 : (code == 2010) ? 3 : 1;
};

// One thread per shell quartet; the blocks stride over the segment, so a
// capped grid (Tune::grid) makes them persistent.  Consecutive quartets walk
// the kets of one bra, so most wavefronts have a single bra: those take the
// wave-uniform path.
template <int LA, int LB, int LC, int LD, typename R, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ClassWaves<LA, LB, LC, LD>::min)))
void k_eri_class(const ClassParams cp)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    // the far one-node launches need no node table
    constexpr int NTAB = MODE == kModeJKFar ? 0 : node_table_lds(S::NR);
    __shared__ double s_tab[NTAB > 0 ? NTAB : 1];
    NodeTable nt{};
    if constexpr (MODE != kModeJKFar) nt = stage_node_table<S::NR>(s_tab);
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long blk = xcd_block();
    const long long lane0 = blk * blockDim.x + (threadIdx.x & ~63u);
    for (long long q = blk * blockDim.x + threadIdx.x, qw = lane0; q < cp.nq;
         q += stride, qw += stride) {
        QuartetIdx ix;
        if constexpr (S::L <= 1) ix = cp.qlist ? locate_list(cp, q, qw) : locate(cp, q, qw);
        else ix = locate(cp, q, qw);
        const int iw = __builtin_amdgcn_readfirstlane(ix.i);
        if (__all(ix.i == iw))
            class_quartet<LA, LB, LC, LD, R, MODE, true>(cp, nt, iw, ix.j, ix.qa);
        else
            class_quartet<LA, LB, LC, LD, R, MODE, false>(cp, nt, ix.i, ix.j, ix.qa);
    }
}

// Schwarz factor Q_ab = sqrt(max_{i in a, j in b} |(ij|ij)|) for one pair
// class, and the same factor Qp of every primitive pair of the shell pair
// taken on its own (the bound of its share of any quartet).
template <int LA, int LB>
__global__ __launch_bounds__(256) void k_schwarz(const int* __restrict__ sa, const int* __restrict__ sb,
                                                 const int* __restrict__ poff, const int* __restrict__ np,
                                                 int n, const PrimPair* __restrict__ prim,
                                                 const double* __restrict__ xyz, double* __restrict__ Q,
                                                 double* __restrict__ Qp)
{
    using S = QuartetShape<LA, LB, LA, LB>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    auto diag_max = [](const double* acc) {
        double m = 0.0;
#pragma unroll
        for (int fa = 0; fa < S::NA; ++fa)
#pragma unroll
            for (int fb = 0; fb < S::NB; ++fb)
                m = fmax(m, fabs(acc[((fa * S::NB + fb) * S::NA + fa) * S::NB + fb]));
        return sqrt(m);
    };
    double acc[S::NINT];
#pragma unroll
    for (int k = 0; k < S::NINT; ++k) acc[k] = 0.0;
    eri_shell_quartet<LA, LB, LA, LB, double>(prim + poff[i], np[i], prim + poff[i], np[i], A, B, A,
                                              B, acc);
    Q[i] = diag_max(acc);
    for (int k = 0; k < np[i]; ++k) {
        const PrimPair* pk = prim + poff[i] + k;
#pragma unroll
        for (int e = 0; e < S::NINT; ++e) acc[e] = 0.0;
        eri_shell_quartet<LA, LB, LA, LB, double>(pk, 1, pk, 1, A, B, A, B, acc);
        Qp[poff[i] + k] = diag_max(acc);
    }
}


// Split classes, pass 1: one wavefront per (64 quartets, bra component
// pair); the component pair index is wave-uniform, so each wavefront runs
// one compile-time instance of eri_bra_component.  Writes raw integral
// planes raw[(fa*NB+fb)*NC*ND + fc*ND + fd][q].
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void k_eri_split(const ClassParams cp, double* __restrict__ raw)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NS = S::NSLICE_ERI, NK = S::NC * S::ND;
    constexpr int NOUT = S::SPLIT_ROWS ? S::NB * NK : NK;  // integrals per lane
    __shared__ double s_tab[node_table_lds(S::NR) > 0 ? node_table_lds(S::NR) : 1];
    const NodeTable nt = stage_node_table<S::NR>(s_tab);
    const long long nwork = ((cp.nq + 63) >> 6) * NS;
    const long long wstride = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; gw < nwork;
         gw += wstride) {
    const int s = __builtin_amdgcn_readfirstlane((int)(gw % NS));
    const long long q = (gw / NS) * 64 + (threadIdx.x & 63);
    if (q >= cp.nq) continue;
    const QuartetIdx ix = locate(cp, q, (gw / NS) * 64);
    const int i = ix.i, j = ix.j;
    const long long qa = ix.qa;
    const int a = cp.bra.sa[i], b = cp.bra.sb[i];
    const double A[3] = {cp.shell_xyz[3 * a], cp.shell_xyz[3 * a + 1], cp.shell_xyz[3 * a + 2]};
    const double B[3] = {cp.shell_xyz[3 * b], cp.shell_xyz[3 * b + 1], cp.shell_xyz[3 * b + 2]};
    // the ket's centres and primitive pairs from the transposed arrays
    const long long kn = cp.kt_n;
    const double C[3] = {cp.kc[j], cp.kc[kn + j], cp.kc[2 * kn + j]};
    const double D[3] = {cp.kc[3 * kn + j], cp.kc[4 * kn + j], cp.kc[5 * kn + j]};
    const KetView kv{cp.kt + j, cp.kt_fs, kn};
    const StridedView kqv{cp.kt + kKetQ * cp.kt_fs + j, kn};
    R acc[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) acc[k] = (R)0;
    // the segment's primitive threshold; in one launch over both mixed
    // segments, a bra's kets past its n64 FP64 ones take the FP32 one
    double tau = cp.tau_prim;
    if (cp.n64 && j - cp.ket0[i] >= cp.n64[i]) tau = cp.tau_prim_lo;
    static_switch<0, NS>(s, [&](auto sc) {
        constexpr bool PS = true;
        if constexpr (S::SPLIT_ROWS) {
            constexpr int FA = decltype(sc)::value;
            eri_bra_row<LA, LB, LC, LD, FA, R, PS>(cp.prim + cp.bra.poff[i], cp.bra.np[i], kv,
                                                   cp.ket.np[j], A, B, C, D, acc, nt,
                                                   cp.primq + cp.bra.poff[i], kqv, tau);
        } else {
            constexpr int FA = decltype(sc)::value / S::NB, FB = decltype(sc)::value % S::NB;
            eri_bra_component<LA, LB, LC, LD, FA, FB, R, PS>(
                cp.prim + cp.bra.poff[i], cp.bra.np[i], kv, cp.ket.np[j], A, B, C, D, acc, nt,
                cp.primq + cp.bra.poff[i], kqv, tau);
        }
    });
    // raw plane of integral (fa fb | fc fd): ((fa*NB + fb)*NK + fc*ND + fd);
    // a row slice s = fa covers planes s*NB*NK .. (s+1)*NB*NK - 1
#pragma unroll
    for (int k = 0; k < NOUT; ++k) raw[(long long)(s * NOUT + k) * cp.qstride + qa] = (double)acc[k];
    }
}

// Split classes, pass 2: one wavefront per (64 quartets, digestion tile),
// or, with `rows` (classes with few quartets, where that leaves most of the
// chip idle), per (64 quartets, tile, tile row).  Same results either way.
template <int LA, int LB, int LC, int LD>
struct TileRows {
    static constexpr int n1(int T)
    {
        using S = QuartetShape<LA, LB, LC, LD>;
        return (T == 0 || T == 2 || T == 3) ? S::NA : (T == 1) ? S::NC : S::NB;
    }
    static constexpr int base(int T) { return T == 0 ? 0 : base(T - 1) + n1(T - 1); }
    static constexpr int total = base(5) + n1(5);
};

// Classes digested by tile rows only: (dd|dd), whose one-tile groups (36
// outputs and a 36-element density block per lane) spill to scratch (304
// bytes per lane) where the rows do not.
template <int LA, int LB, int LC, int LD>
struct DigestRowsOnly {
    static constexpr bool value = LA == 2 && LB == 2 && LC == 2 && LD == 2;
};

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void k_digest_split(const ClassParams cp,
                                                      const double* __restrict__ raw, int rows)
{
    using TR = TileRows<LA, LB, LC, LD>;
    using TG = TileGroups<LA, LB, LC, LD>;
    if constexpr (DigestRowsOnly<LA, LB, LC, LD>::value) rows = 1;
    const int per = rows ? TR::total : TG::count;
    const long long nwork = ((cp.nq + 63) >> 6) * per;
    const long long wstride = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long gw = ((long long)xcd_block() * blockDim.x + threadIdx.x) >> 6; gw < nwork;
         gw += wstride) {
    const int w = __builtin_amdgcn_readfirstlane((int)(gw % per));
    int t = w, r = 0;
    if (rows) {
        t = 0;
        while (t < 5 && w >= TR::base(t + 1)) ++t;
        r = w - TR::base(t);
    }
    const long long q = (gw / per) * 64 + (threadIdx.x & 63);
    if (q >= cp.nq) continue;
    const QuartetIdx ix = locate(cp, q, (gw / per) * 64);
    const int i = ix.i, j = ix.j;
    const long long qa = ix.qa;
    const int a = cp.bra.sa[i], b = cp.bra.sb[i];
    const int c = cp.ket.sa[j], d = cp.ket.sb[j];
    const double fac = ((a == b) ? 0.5 : 1.0) * ((c == d) ? 0.5 : 1.0) *
                       ((cp.same && i == j) ? 0.5 : 1.0);
    const QuartetMap qm = quartet_map(cp.map, a, b, c, d);
    if (rows) {
        static_switch<0, 6>(t, [&](auto tc) {
            constexpr int T = decltype(tc)::value;
            using G = TileGeom<LA, LB, LC, LD, T>;
            double* dst = cp.tiles + cp.pos[T * cp.qstride + qa];
            double Pm[G::M1][G::M2];
            load_tile_density<LA, LB, LC, LD, T>(cp.P, qm, Pm);
            double row[G::N2];
            digest_tile_row<LA, LB, LC, LD, T>(raw + qa, cp.qstride, fac, Pm, r, row);
#pragma unroll
            for (int e = 0; e < G::N2; ++e) dst[r * G::N2 + e] = row[e];
        });
    } else if constexpr (!DigestRowsOnly<LA, LB, LC, LD>::value) {
        // w = the tile group
        static_switch<0, TG::count>(w, [&](auto gc) {
            constexpr int g = decltype(gc)::value;
            double* dst[6];
#pragma unroll
            for (int T = 0; T < 6; ++T) dst[T] = cp.tiles + cp.pos[T * cp.qstride + qa];
            digest_tile_group<LA, LB, LC, LD, TG::begin(g), TG::end(g)>(raw + qa, cp.qstride, fac,
                                                                        cp.P, qm, dst);
        });
    }
    }
}

// Nuclear gradient, two-electron part: one wavefront per (64 quartets, bra
// component pair).  Each thread contracts d(fa fb|fc fd)/d(A,B,C) with the
// density factor fac * [4 P_ab P_cd - P_ac P_bd - P_ad P_bc] (the 8 images of
// 1/2 sum Gamma_pqrs (pq|rs)^x with Gamma = P_pq P_rs - P_pr P_qs / 2) and
// writes its 9 partial derivatives to gout[(s*9 + k) * qstride + q].
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void k_grad_split(const ClassParams cp, double* __restrict__ gout)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    constexpr int NS = S::NSLICE_GRAD, NK = S::NC * S::ND;
    // derivative integrals need one node more than the integrals (L + 1)
    constexpr int NRG = (S::L + 1) / 2 + 1;
    __shared__ double s_tab[node_table_lds(NRG) > 0 ? node_table_lds(NRG) : 1];
    const NodeTable nt = stage_node_table<NRG>(s_tab);  // every thread, before any exit
    const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int s = __builtin_amdgcn_readfirstlane((int)(gw % NS));
    const long long q = (gw / NS) * 64 + (threadIdx.x & 63);
    if (q >= cp.nq) return;
    const QuartetIdx ix = locate(cp, q, (gw / NS) * 64);
    const int i = ix.i, j = ix.j;
    const long long qa = ix.qa;
    const int a = cp.bra.sa[i], b = cp.bra.sb[i];
    const int c = cp.ket.sa[j], d = cp.ket.sb[j];
    const double A[3] = {cp.shell_xyz[3 * a], cp.shell_xyz[3 * a + 1], cp.shell_xyz[3 * a + 2]};
    const double B[3] = {cp.shell_xyz[3 * b], cp.shell_xyz[3 * b + 1], cp.shell_xyz[3 * b + 2]};
    const double C[3] = {cp.shell_xyz[3 * c], cp.shell_xyz[3 * c + 1], cp.shell_xyz[3 * c + 2]};
    const double D[3] = {cp.shell_xyz[3 * d], cp.shell_xyz[3 * d + 1], cp.shell_xyz[3 * d + 2]};
    const double fac = ((a == b) ? 0.5 : 1.0) * ((c == d) ? 0.5 : 1.0) *
                       ((cp.same && i == j) ? 0.5 : 1.0);
    const QuartetMap qm = quartet_map(cp.map, a, b, c, d);
    const double* P = cp.P;
    const long long n = qm.ld;
    // the segment's primitive threshold; in one launch over both mixed
    // segments, a bra's kets past its n64 FP64 ones take the FP32 one
    double tau = cp.tau_prim;
    if (cp.n64 && j - cp.ket0[i] >= cp.n64[i]) tau = cp.tau_prim_lo;
    double g[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = 0.0;
    if constexpr (S::GRAD_ALL_A) {
        // every component of A in this thread (B is an s shell): one slice
        double Df[S::NA * NK];
        const double* Pb = P + qm.row[1];   // row 0 (the s function) of shell B
#pragma unroll
        for (int fa = 0; fa < S::NA; ++fa) {
            const double* Pa = P + qm.row[0] + fa * n;
            const double pab = Pa[qm.col[1]];
#pragma unroll
            for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
                for (int fd = 0; fd < S::ND; ++fd)
                    Df[fa * NK + fc * S::ND + fd] =
                        fac * (4.0 * pab * P[qm.row[2] + fc * n + qm.col[3] + fd] -
                               Pa[qm.col[2] + fc] * Pb[qm.col[3] + fd] -
                               Pa[qm.col[3] + fd] * Pb[qm.col[2] + fc]);
        }
        grad_all_a<LA, LB, LC, LD, R>(cp.prim + cp.bra.poff[i], cp.bra.np[i], cp.prim + cp.ket.poff[j],
                                      cp.ket.np[j], A, B, C, D, Df, g, nt,
                                      cp.primq + cp.bra.poff[i], cp.primq + cp.ket.poff[j], tau);
    } else {
    static_switch<0, NS>(s, [&](auto sc) {
        constexpr int FA = decltype(sc)::value / S::NB, FB = decltype(sc)::value % S::NB;
        const double* Pa = P + qm.row[0] + FA * n;   // row fa of P
        const double* Pb = P + qm.row[1] + FB * n;   // row fb of P
        const double pab = Pa[qm.col[1] + FB];
        double Df[NK];
#pragma unroll
        for (int fc = 0; fc < S::NC; ++fc)
#pragma unroll
            for (int fd = 0; fd < S::ND; ++fd)
                Df[fc * S::ND + fd] =
                    fac * (4.0 * pab * P[qm.row[2] + fc * n + qm.col[3] + fd] -
                           Pa[qm.col[2] + fc] * Pb[qm.col[3] + fd] -
                           Pa[qm.col[3] + fd] * Pb[qm.col[2] + fc]);
        grad_bra_component<LA, LB, LC, LD, FA, FB, R>(cp.prim + cp.bra.poff[i], cp.bra.np[i],
                                                      cp.prim + cp.ket.poff[j], cp.ket.np[j], A,
                                                      B, C, D, Df, g, nt,
                                                      cp.primq + cp.bra.poff[i],
                                                      cp.primq + cp.ket.poff[j], tau);
    });
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) gout[(long long)(s * 9 + k) * cp.qstride + qa] = g[k];
}

template <int LA, int LB, int LC, int LD>
static int launch_grad_t(int fp32, const ClassParams& cp, double* gout, hipStream_t s)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    if (cp.nq <= 0) return kOk;
    const long long waves = (cp.nq + 63) / 64;
    const long long nb = (waves * S::NSLICE_GRAD + 3) / 4;
    if (nb > 0x7fffffffLL) {
        set_error(kEUnsupported, "quartet segment too large (%lld)", cp.nq);
        return kEUnsupported;
    }
    if (fp32)
        hipLaunchKernelGGL((k_grad_split<LA, LB, LC, LD, float>), dim3((unsigned)nb), dim3(256), 0, s,
                           cp, gout);
    else
        hipLaunchKernelGGL((k_grad_split<LA, LB, LC, LD, double>), dim3((unsigned)nb), dim3(256), 0,
                           s, cp, gout);
    SF_HIP(hipGetLastError());
    return kOk;
}

// Blocks for `items` work items of `per_block` each, capped by the tuned
// persistent grid.  Fails when the launch would exceed the grid limit.
static int grid_blocks(long long items, long long per_block, const Tune& tn, unsigned* out)
{
    long long nb = (items + per_block - 1) / per_block;
    if (tn.grid > 0 && nb > tn.grid) nb = tn.grid;
    if (nb > 0x7fffffffLL) {
        set_error(kEUnsupported, "quartet segment too large (%lld work items)", items);
        return kEUnsupported;
    }
    *out = (unsigned)(nb > 0 ? nb : 1);
    return kOk;
}

static int check_block(const Tune& tn)
{
    if (tn.block < 64 || tn.block > 256 || tn.block % 64 != 0 || tn.grid < 0) {
        set_error(kEInval, "launch shape block=%d grid=%d (block 64..256, multiple of 64)", tn.block,
                  tn.grid);
        return kEInval;
    }
    return kOk;
}

template <int LA, int LB, int LC, int LD>
static int launch_split(int fp32, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    using S = QuartetShape<LA, LB, LC, LD>;
    const long long waves = (cp.nq + 63) / 64;
    const int wpb = tn.block / 64;
    unsigned nb1 = 0, nb2 = 0;
    SF_TRY_K(grid_blocks(waves * S::NSLICE_ERI, wpb, tn, &nb1));
    // digestion by tile rows when by tiles would give fewer than 1536
    // wavefronts (6 per CU): measured faster below, slower above
    const int rows = (DigestRowsOnly<LA, LB, LC, LD>::value || waves * 6 < 1536) ? 1 : 0;
    SF_TRY_K(grid_blocks(waves * (rows ? TileRows<LA, LB, LC, LD>::total
                                       : TileGroups<LA, LB, LC, LD>::count),
                         wpb, tn, &nb2));
    double* raw = (mode == kModeRaw) ? cp.out : cp.raw;
    if (!raw) {
        set_error(kEState, "split class without a raw buffer");
        return kEState;
    }
    if (mode != kModeDigestOnly) {
        if (fp32 && mode != kModeRaw)
            hipLaunchKernelGGL((k_eri_split<LA, LB, LC, LD, float>), dim3(nb1), dim3(tn.block), 0,
                               s, cp, raw);
        else
            hipLaunchKernelGGL((k_eri_split<LA, LB, LC, LD, double>), dim3(nb1), dim3(tn.block), 0,
                               s, cp, raw);
        SF_HIP(hipGetLastError());
    }
    if (mode == kModeJK || mode == kModeDigestOnly) {
        hipLaunchKernelGGL((k_digest_split<LA, LB, LC, LD>), dim3(nb2), dim3(tn.block), 0, s, cp,
                           (const double*)raw, rows);
        SF_HIP(hipGetLastError());
    }
    return kOk;
}

template <int LA, int LB, int LC, int LD>
static int launch_class(int fp32, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    if (cp.nq <= 0) return kOk;
    SF_TRY_K(check_block(tn));
    if constexpr (QuartetShape<LA, LB, LC, LD>::SPLIT) {
        return launch_split<LA, LB, LC, LD>(fp32, mode, cp, tn, s);
    } else {
        if (mode == kModeEriOnly || mode == kModeDigestOnly) {
            set_error(kEInval, "separate integral and digestion launches are for split classes");
            return kEInval;
        }
        unsigned nb = 0;
        SF_TRY_K(grid_blocks(cp.nq, tn.block, tn, &nb));
        dim3 grid(nb), block(tn.block);
        if constexpr (QuartetShape<LA, LB, LC, LD>::L <= 1) {
            if (mode == kModeJK && cp.far) {
                if (fp32)
                    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, float, kModeJKFar>), grid, block,
                                       0, s, cp);
                else
                    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, double, kModeJKFar>), grid,
                                       block, 0, s, cp);
                SF_HIP(hipGetLastError());
                return kOk;
            }
        }
        if (mode == kModeRaw)
            hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, double, kModeRaw>), grid, block, 0, s, cp);
        else if (fp32)
            hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, float, kModeJK>), grid, block, 0, s, cp);
        else
            hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, double, kModeJK>), grid, block, 0, s, cp);
        SF_HIP(hipGetLastError());
        return kOk;
    }
}

// McMurchie-Davidson engine (eri_mmd.hpp): one wavefront per shell
// quartet, blocks stride over the segment; raw integral planes.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_eri_mmd(const ClassParams cp, double* __restrict__ out)
{
    using M = MmdShape<LA, LB, LC, LD>;
    __shared__ double sR[2][herm_n(M::L)];
    __shared__ double sEab[3][M::EAB];
    __shared__ double sEcd[3][M::ECD];
    __shared__ double sT[M::TR * M::TC];
    for (long long q = blockIdx.x; q < cp.nq; q += gridDim.x) {
        const QuartetIdx ix = locate(cp, q, q);
        const int i = ix.i, j = ix.j;
        const int a = cp.bra.sa[i], b = cp.bra.sb[i];
        const int c = cp.ket.sa[j], d = cp.ket.sb[j];
        const double A[3] = {cp.shell_xyz[3 * a], cp.shell_xyz[3 * a + 1], cp.shell_xyz[3 * a + 2]};
        const double B[3] = {cp.shell_xyz[3 * b], cp.shell_xyz[3 * b + 1], cp.shell_xyz[3 * b + 2]};
        const double C[3] = {cp.shell_xyz[3 * c], cp.shell_xyz[3 * c + 1], cp.shell_xyz[3 * c + 2]};
        const double D[3] = {cp.shell_xyz[3 * d], cp.shell_xyz[3 * d + 1], cp.shell_xyz[3 * d + 2]};
        mmd_shell_quartet<LA, LB, LC, LD>(cp.prim + cp.bra.poff[i], cp.bra.np[i],
                                          cp.prim + cp.ket.poff[j], cp.ket.np[j], A, B, C, D, out,
                                          cp.qstride, ix.qa, sR, sEab, sEcd, sT);
    }
}

// MMD integrals into raw planes, then (J/K mode) the raw-plane digestion.
template <int LA, int LB, int LC, int LD>
static int launch_mmd_t(int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    if (cp.nq <= 0) return kOk;
    SF_TRY_K(check_block(tn));
    double* raw = (mode == kModeRaw) ? cp.out : cp.raw;
    if (!raw) {
        set_error(kEState, "MMD engine without a raw integral buffer");
        return kEState;
    }
    long long nb = cp.nq;
    const long long cap = tn.grid > 0 ? tn.grid : 16384;
    if (nb > cap) nb = cap;
    hipLaunchKernelGGL((k_eri_mmd<LA, LB, LC, LD>), dim3((unsigned)nb), dim3(64), 0, s, cp, raw);
    SF_HIP(hipGetLastError());
    if (mode == kModeJK) {
        unsigned nb2 = 0;
        const long long waves = (cp.nq + 63) / 64;
        const int rows = (DigestRowsOnly<LA, LB, LC, LD>::value || waves * 6 < 1536) ? 1 : 0;
        SF_TRY_K(grid_blocks(waves * (rows ? TileRows<LA, LB, LC, LD>::total
                                           : TileGroups<LA, LB, LC, LD>::count),
                             4, Tune(), &nb2));
        hipLaunchKernelGGL((k_digest_split<LA, LB, LC, LD>), dim3(nb2), dim3(256), 0, s, cp,
                           (const double*)raw, rows);
        SF_HIP(hipGetLastError());
    }
    return kOk;
}

template <int LA, int LB, int LC, int LD>
static int launch_schwarz_t(const int* sa, const int* sb, const int* poff, const int* np, int n,
                            const PrimPair* prim, const double* xyz, double* Q, double* Qp,
                            hipStream_t s)
{
    if constexpr (LA == LC && LB == LD) {
        if (n <= 0) return kOk;
        hipLaunchKernelGGL((k_schwarz<LA, LB>), dim3((n + 127) / 128), dim3(128), 0, s, sa, sb, poff,
                           np, n, prim, xyz, Q, Qp);
        SF_HIP(hipGetLastError());
        return kOk;
    } else {
        set_error(kEUnsupported, "no Schwarz kernel in an off-diagonal class");
        return kEUnsupported;
    }
}

// With SF_PART defined, a class unit holds one part of its kernels (0: J/K
// and Schwarz, 1: gradient, 2: McMurchie-Davidson), so the heavy classes
// compile in parallel; without it, all of them.
#if !defined(SF_PART) || SF_PART == 0
#define SF_INST_ERI(id, a, b, c, d)                                                             \
    int launch_class_##id(int fp32, int mode, const ClassParams& cp, const Tune& tn,           \
                          hipStream_t s)                                                       \
    {                                                                                          \
        return launch_class<a, b, c, d>(fp32, mode, cp, tn, s);                                \
    }                                                                                          \
    int launch_schwarz_##id(const int* sa, const int* sb, const int* poff, const int* np, int n, \
                            const PrimPair* prim, const double* xyz, double* Q, double* Qp,     \
                            hipStream_t s)                                                     \
    {                                                                                          \
        return launch_schwarz_t<a, b, c, d>(sa, sb, poff, np, n, prim, xyz, Q, Qp, s);         \
    }
#else
#define SF_INST_ERI(id, a, b, c, d)
#endif
#if !defined(SF_PART) || SF_PART == 1
#define SF_INST_GRAD(id, a, b, c, d)                                                            \
    int launch_grad_##id(int fp32, const ClassParams& cp, double* gout, hipStream_t s)         \
    {                                                                                          \
        return launch_grad_t<a, b, c, d>(fp32, cp, gout, s);                                   \
    }
#else
#define SF_INST_GRAD(id, a, b, c, d)
#endif
#if !defined(SF_PART) || SF_PART == 2
#define SF_INST_MMD(id, a, b, c, d)                                                             \
    int launch_mmd_##id(int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)        \
    {                                                                                          \
        return launch_mmd_t<a, b, c, d>(mode, cp, tn, s);                                      \
    }
#else
#define SF_INST_MMD(id, a, b, c, d)
#endif
#define SF_INST(id, a, b, c, d) \
    SF_INST_ERI(id, a, b, c, d) SF_INST_GRAD(id, a, b, c, d) SF_INST_MMD(id, a, b, c, d)
#define SF_PICK(id, a, b, c, d) SF_PICK_##id(id, a, b, c, d)
#if SF_CLASS == 0
#define SF_PICK_0(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_0(id, a, b, c, d)
#endif
#if SF_CLASS == 1
#define SF_PICK_1(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_1(id, a, b, c, d)
#endif
#if SF_CLASS == 2
#define SF_PICK_2(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_2(id, a, b, c, d)
#endif
#if SF_CLASS == 3
#define SF_PICK_3(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_3(id, a, b, c, d)
#endif
#if SF_CLASS == 4
#define SF_PICK_4(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_4(id, a, b, c, d)
#endif
#if SF_CLASS == 5
#define SF_PICK_5(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_5(id, a, b, c, d)
#endif
#if SF_CLASS == 6
#define SF_PICK_6(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_6(id, a, b, c, d)
#endif
#if SF_CLASS == 7
#define SF_PICK_7(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_7(id, a, b, c, d)
#endif
#if SF_CLASS == 8
#define SF_PICK_8(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_8(id, a, b, c, d)
#endif
#if SF_CLASS == 9
#define SF_PICK_9(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_9(id, a, b, c, d)
#endif
#if SF_CLASS == 10
#define SF_PICK_10(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_10(id, a, b, c, d)
#endif
#if SF_CLASS == 11
#define SF_PICK_11(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_11(id, a, b, c, d)
#endif
#if SF_CLASS == 12
#define SF_PICK_12(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_12(id, a, b, c, d)
#endif
#if SF_CLASS == 13
#define SF_PICK_13(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_13(id, a, b, c, d)
#endif
#if SF_CLASS == 14
#define SF_PICK_14(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_14(id, a, b, c, d)
#endif
#if SF_CLASS == 15
#define SF_PICK_15(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_15(id, a, b, c, d)
#endif
#if SF_CLASS == 16
#define SF_PICK_16(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_16(id, a, b, c, d)
#endif
#if SF_CLASS == 17
#define SF_PICK_17(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_17(id, a, b, c, d)
#endif
#if SF_CLASS == 18
#define SF_PICK_18(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_18(id, a, b, c, d)
#endif
#if SF_CLASS == 19
#define SF_PICK_19(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_19(id, a, b, c, d)
#endif
#if SF_CLASS == 20
#define SF_PICK_20(id, a, b, c, d) SF_INST(id, a, b, c, d)
#else
#define SF_PICK_20(id, a, b, c, d)
#endif

SF_FOR_EACH_CLASS(SF_PICK)

// An f-shell class unit: the J/K / dense-export launcher of one class.
#ifdef SF_FCLASS
#define SF_FNAME2(id) launch_class_##id
#define SF_FNAME(id) SF_FNAME2(id)
int SF_FNAME(SF_FCLASS)(int fp32, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    static_assert(quartet_class(pair_class(SF_FLA, SF_FLB), pair_class(SF_FLC, SF_FLD)) ==
                      SF_FCLASS, "f class id and angular momenta disagree");
    return launch_class<SF_FLA, SF_FLB, SF_FLC, SF_FLD>(fp32, mode, cp, tn, s);
}
#endif

#else  // dispatch tables

static_assert(quartet_class(pair_class(2, 1), pair_class(1, 1)) == 12, "class numbering");

// Primitive quartets the Rys class kernels evaluate in one segment: the
// same ket-outer / bra-inner loops and threshold breaks as the screened
// primitive loops of eri_shell_quartet, counted instead of computed.  One
// partial sum per block (no atomics).
__global__ __launch_bounds__(256) void k_count_prims(const ClassParams cp,
                                                     unsigned long long* __restrict__ part)
{
    __shared__ unsigned long long s_sum[256];
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long lane0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    unsigned long long cnt = 0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x, qw = lane0; q < cp.nq;
         q += stride, qw += stride) {
        const QuartetIdx ix = cp.qlist ? locate_list(cp, q, qw) : locate(cp, q, qw);
        const double* bq = cp.primq + cp.bra.poff[ix.i];
        const double* kq = cp.primq + cp.ket.poff[ix.j];
        const int nb = cp.bra.np[ix.i], nk = cp.ket.np[ix.j];
        for (int ik = 0; ik < nk; ++ik) {
            const double qk = kq[ik];
            if (nb == 0 || qk * bq[0] < cp.tau_prim) break;
            int ib = 0;
            while (ib < nb && !(bq[ib] * qk < cp.tau_prim)) ++ib;
            cnt += (unsigned long long)ib;
        }
    }
    s_sum[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_sum[0];
}

int launch_count_prims(const ClassParams& cp, int nblock, unsigned long long* part, hipStream_t s)
{
    if (nblock <= 0 || nblock > 65535) {
        set_error(kEInval, "launch_count_prims: %d blocks", nblock);
        return kEInval;
    }
    hipLaunchKernelGGL(k_count_prims, dim3(nblock), dim3(256), 0, s, cp, part);
    SF_HIP(hipGetLastError());
    return kOk;
}

// (la, lb, lc, ld) of quartet class c.
static GenClass class_shape(int c)
{
    const int pcb = pair_of_class(c), pck = c - pcb * (pcb + 1) / 2;
    return gen_class(pair_la(pcb), pair_lb(pcb), pair_la(pck), pair_lb(pck));
}

int class_generic(int qcls)
{
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return 0;
        SF_FOR_EACH_CLASS(X)
        SF_FOR_EACH_FCLASS(X)
#undef X
    }
    return (qcls >= kNumClassSPD && qcls < kNumQuartetClass) ? 1 : 0;
}

int launch_eri_class(int qcls, int fp32, int mode, const ClassParams& cp, const Tune& tn,
                     hipStream_t s)
{
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return launch_class_##id(fp32, mode, cp, tn, s);
        SF_FOR_EACH_CLASS(X)
        SF_FOR_EACH_FCLASS(X)
#undef X
    }
    // the larger f-shell classes: the run-time-L kernels, FP64 arithmetic
    if (class_generic(qcls)) return launch_eri_generic(qcls, mode, cp, tn, s);
    set_error(kEUnsupported, "quartet class %d", qcls);
    return kEUnsupported;
}

int launch_eri_mmd(int qcls, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    if (qcls >= kNumClassSPD && qcls < kNumQuartetClass) {
        set_error(kEUnsupported, "quartet class %d: the McMurchie-Davidson engine covers s, p, d "
                  "shells only", qcls);
        return kEUnsupported;
    }
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return launch_mmd_##id(mode, cp, tn, s);
        SF_FOR_EACH_CLASS(X)
#undef X
    }
    set_error(kEUnsupported, "quartet class %d", qcls);
    return kEUnsupported;
}

int class_ntile(int qcls)
{
    if (qcls >= kNumClassSPD && qcls < kNumQuartetClass) return class_shape(qcls).ntile;
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return QuartetShape<a, b, c, d>::NTILE;
        SF_FOR_EACH_CLASS(X)
#undef X
    }
    return 0;
}

int launch_grad_class(int qcls, int fp32, const ClassParams& cp, double* gout, hipStream_t s)
{
    if (qcls >= kNumClassSPD && qcls < kNumQuartetClass) {
        set_error(kEUnsupported, "quartet class %d: nuclear gradients cover s, p, d shells only",
                  qcls);
        return kEUnsupported;
    }
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return launch_grad_##id(fp32, cp, gout, s);
        SF_FOR_EACH_CLASS(X)
#undef X
    }
    set_error(kEUnsupported, "quartet class %d", qcls);
    return kEUnsupported;
}

int class_nslice(int qcls)
{
    if (qcls >= kNumClassSPD) return 0;
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return QuartetShape<a, b, c, d>::NSLICE_GRAD;
        SF_FOR_EACH_CLASS(X)
#undef X
    }
    return 0;
}

int class_split(int qcls)
{
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return QuartetShape<a, b, c, d>::SPLIT ? 1 : 0;
        SF_FOR_EACH_CLASS(X)
        SF_FOR_EACH_FCLASS(X)
#undef X
    }
    // the run-time-L classes write raw planes and digest them after
    if (class_generic(qcls)) return 1;
    return 0;
}

int class_nint(int qcls)
{
    if (qcls >= kNumClassSPD && qcls < kNumQuartetClass) return class_shape(qcls).nint;
    switch (qcls) {
#define X(id, a, b, c, d) \
    case id: return QuartetShape<a, b, c, d>::NINT;
        SF_FOR_EACH_CLASS(X)
#undef X
    }
    return 0;
}

void class_tile_offsets(int qcls, int off[6])
{
    if (qcls >= kNumClassSPD && qcls < kNumQuartetClass) {
        const GenClass g = class_shape(qcls);
        for (int t = 0; t < 6; ++t) off[t] = g.toff[t];
        return;
    }
    switch (qcls) {
#define X(id, a, b, c, d)                                        \
    case id: {                                                   \
        using S = QuartetShape<a, b, c, d>;                      \
        off[0] = S::T_AB; off[1] = S::T_CD; off[2] = S::T_AC;    \
        off[3] = S::T_AD; off[4] = S::T_BC; off[5] = S::T_BD;    \
        return;                                                  \
    }
        SF_FOR_EACH_CLASS(X)
#undef X
    }
}

int launch_schwarz(int pc, const int* sa, const int* sb, const int* poff, const int* np, int n,
                   const PrimPair* prim, const double* shell_xyz, double* Q, double* Qp,
                   hipStream_t s)
{
    // the diagonal quartet class (pc|pc) holds the pair class's Schwarz kernel
    static const SchwarzLaunch diag[kNumPairSPD] = {launch_schwarz_0, launch_schwarz_2,
                                                         launch_schwarz_5, launch_schwarz_9,
                                                         launch_schwarz_14, launch_schwarz_20};
    if (pc < 0 || pc >= kNumPairClass) {
        set_error(kEUnsupported, "pair class %d", pc);
        return kEUnsupported;
    }
    if (pc >= kNumPairSPD)
        return launch_schwarz_generic(pc, sa, sb, poff, np, n, prim, shell_xyz, Q, Qp, s);
    return diag[pc](sa, sb, poff, np, n, prim, shell_xyz, Q, Qp, s);
}

#endif  // SF_CLASS

}  // namespace sf
