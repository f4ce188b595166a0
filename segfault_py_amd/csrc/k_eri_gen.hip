// k_eri_gen.hip -- quartet classes with an f shell (L = 3): integrals,
// digestion and Schwarz factors with the angular momenta as run-time values.
//
// The s/p/d classes and the smaller f classes have per-class kernels with
// compile-time bra components (k_eri.hip); their compile time grows steeply
// with the class size, so the 15 largest f classes run the kernels here:
//   * k_eri_gen<LA, LB, LC, LD>: one wavefront per (64 quartets, bra
//     component pair), the pair wave-uniform and a run-time value
//     (eri_bra_runtime), so one compiled instance per class: each lane
//     contracts its own shell quartet over the primitive pairs (ket outer,
//     bra inner, the screened loops of the class kernels) and writes the
//     raw integral planes of all ket components of that bra pair.
//   * k_digest_gen: one wavefront per (64 quartets, digestion tile element):
//     each lane sums that element of its quartet's tile from the raw planes
//     and the density, and stores it in the tile's destination-sorted slot.
//   * k_schwarz_gen: one thread per shell pair: Q = sqrt(max |(ab|ab)|) over
//     the pair's component pairs, and the same for each primitive pair alone.
// Nodes: floor(L/2) + 1 <= 7 (L <= 12), from the 1..7-node Chebyshev tables.
//
// Reference: rys.eri_rys / eri_3d_rys / recur / shift (src/hf/core/
// core.f90:815-1002), which take any angular momentum; it always sums 5
// nodes, which is exact only up to L = 9, so for f-shell quartets with
// L >= 10 this engine (exact) and the reference differ by the reference's
// quadrature error.  The J/K digestion replaces hf.py:300.
#include "sf_internal.hpp"

// The f-shell classes without compiled per-class kernels in k_eri.hip (more
// than 540 integrals per quartet, or more than 54 per component of A).
#define SF_FOR_EACH_GENCLASS(X)                                                                 \
    X(33, 3, 1, 2, 2) X(35, 3, 1, 3, 1) X(40, 3, 2, 2, 1) X(41, 3, 2, 2, 2) X(42, 3, 2, 3, 0)   \
    X(43, 3, 2, 3, 1) X(44, 3, 2, 3, 2) X(47, 3, 3, 1, 1) X(48, 3, 3, 2, 0) X(49, 3, 3, 2, 1)   \
    X(50, 3, 3, 2, 2) X(51, 3, 3, 3, 0) X(52, 3, 3, 3, 1) X(53, 3, 3, 3, 2) X(54, 3, 3, 3, 3)

namespace sf {

// Raw integral planes of one precision segment (or the whole class): one
// wavefront per (64 quartets, bra component pair s = fa * NB + fb), the pair
// wave-uniform; dst[(s * NC * ND + k) * qstride + qa] for ket component k
// of the quartet at class-wide index qa.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void k_eri_gen(const ClassParams cp, double* __restrict__ dst)
{
    constexpr int NB = cart_n(LB), NS = cart_n(LA) * NB, NK = cart_n(LC) * cart_n(LD);
    const long long nwork = ((cp.nq + 63) >> 6) * NS;
    const long long wstride = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; gw < nwork;
         gw += wstride) {
        const int s = __builtin_amdgcn_readfirstlane((int)(gw % NS));
        const long long qw = (gw / NS) * 64;
        const long long q = qw + (threadIdx.x & 63);
        if (q >= cp.nq) continue;
        const QuartetIdx ix = locate(cp, q, qw);
        const int i = ix.i, j = ix.j;
        const int a = cp.bra.sa[i], b = cp.bra.sb[i], c = cp.ket.sa[j], d = cp.ket.sb[j];
        const double A[3] = {cp.shell_xyz[3 * a], cp.shell_xyz[3 * a + 1], cp.shell_xyz[3 * a + 2]};
        const double B[3] = {cp.shell_xyz[3 * b], cp.shell_xyz[3 * b + 1], cp.shell_xyz[3 * b + 2]};
        const double C[3] = {cp.shell_xyz[3 * c], cp.shell_xyz[3 * c + 1], cp.shell_xyz[3 * c + 2]};
        const double D[3] = {cp.shell_xyz[3 * d], cp.shell_xyz[3 * d + 1], cp.shell_xyz[3 * d + 2]};
        // one launch over both mixed segments: kets past the bra's n64 FP64
        // ones take the FP32 segment's primitive threshold
        double tau = cp.tau_prim;
        if (cp.n64 && j - cp.ket0[i] >= cp.n64[i]) tau = cp.tau_prim_lo;
        const int bpo = cp.bra.poff[i], kpo = cp.ket.poff[j];
        double acc[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) acc[k] = 0.0;
        eri_bra_runtime<LA, LB, LC, LD>(s / NB, s % NB, cp.prim + bpo, cp.bra.np[i], cp.prim + kpo,
                                        cp.ket.np[j], A, B, C, D, acc, cp.primq + bpo,
                                        cp.primq + kpo, tau);
#pragma unroll
        for (int k = 0; k < NK; ++k) dst[(long long)(s * NK + k) * cp.qstride + ix.qa] = acc[k];
    }
}

// One wavefront per (64 quartets, tile element): gen_tile_element of each
// lane's quartet, stored in the tile's destination-sorted slot.
__global__ __launch_bounds__(256) void k_digest_gen(const ClassParams cp, const GenClass g,
                                                    const double* __restrict__ raw)
{
    const long long nwork = ((cp.nq + 63) >> 6) * g.ntile;
    const long long wstride = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long gw = ((long long)xcd_block() * blockDim.x + threadIdx.x) >> 6; gw < nwork;
         gw += wstride) {
        const int w = __builtin_amdgcn_readfirstlane((int)(gw % g.ntile));
        int T = 0;
        while (T < 5 && w >= g.toff[T + 1]) ++T;
        const int e = w - g.toff[T];
        const long long qw = (gw / g.ntile) * 64;
        const long long q = qw + (threadIdx.x & 63);
        if (q >= cp.nq) continue;
        const QuartetIdx ix = locate(cp, q, qw);
        const int i = ix.i, j = ix.j;
        const long long qa = ix.qa;
        const int sh[4] = {cp.bra.sa[i], cp.bra.sb[i], cp.ket.sa[j], cp.ket.sb[j]};
        const double fac = ((sh[0] == sh[1]) ? 0.5 : 1.0) * ((sh[2] == sh[3]) ? 0.5 : 1.0) *
                           ((cp.same && i == j) ? 0.5 : 1.0);
        QuartetMap qm;
        for (int k = 0; k < 4; ++k) {
            qm.row[k] = cp.map.row[sh[k]];
            qm.col[k] = cp.map.col[sh[k]];
        }
        qm.ld = cp.map.ld[sh[0]];
        const double t = gen_tile_element(g, T, e, raw + qa, cp.qstride, fac, cp.P, qm);
        cp.tiles[cp.pos[T * cp.qstride + qa] + e] = t;
    }
}

// Schwarz factors of one f-shell pair class: Q[i] for shell pair i and
// Qp[poff[i] + k] for each of its primitive pairs alone (as k_schwarz).
__global__ __launch_bounds__(64) void k_schwarz_gen(const int* __restrict__ sa,
                                                    const int* __restrict__ sb,
                                                    const int* __restrict__ poff,
                                                    const int* __restrict__ np, int n,
                                                    const PrimPair* __restrict__ prim,
                                                    const double* __restrict__ xyz, GenClass g,
                                                    double* __restrict__ Q,
                                                    double* __restrict__ Qp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    const int na = g.n[0], nbc = g.n[1];
    double qmax = 0.0;
    for (int kp = -1; kp < np[i]; ++kp) {
        // kp = -1: the contracted pair; kp >= 0: primitive pair kp alone
        const PrimPair* pp = prim + poff[i] + (kp < 0 ? 0 : kp);
        const int npp = kp < 0 ? np[i] : 1;
        double m = 0.0;
        for (int fa = 0; fa < na; ++fa)
            for (int fb = 0; fb < nbc; ++fb) {
                int lc[12];
                for (int e = 0; e < 3; ++e) {
                    lc[e] = lc[6 + e] = cart_l(g.l[0], fa, e);
                    lc[3 + e] = lc[9 + e] = cart_l(g.l[1], fb, e);
                }
                double v = 0.0;
                static_switch<1, 8>(g.nr, [&](auto nc) {
                    constexpr int NR = decltype(nc)::value;
                    v = gen_component<NR>(lc, A, B, A, B, pp, npp, pp, npp, nullptr, nullptr, 0.0);
                });
                m = fmax(m, fabs(v));
            }
        if (kp < 0) qmax = sqrt(m);
        else Qp[poff[i] + kp] = sqrt(m);
    }
    Q[i] = qmax;
}

static unsigned gen_blocks(long long waves, const Tune& tn)
{
    const long long per = tn.block / 64;
    long long nb = (waves + per - 1) / per;
    const long long cap = tn.grid > 0 ? tn.grid : (1LL << 20);
    if (nb > cap) nb = cap;
    return (unsigned)(nb > 0 ? nb : 1);
}

int launch_eri_generic(int qcls, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s)
{
    if (!class_generic(qcls)) {
        set_error(kEInval, "quartet class %d has compiled kernels", qcls);
        return kEInval;
    }
    const int pcb = pair_of_class(qcls), pck = qcls - pcb * (pcb + 1) / 2;
    const GenClass g = gen_class(pair_la(pcb), pair_lb(pcb), pair_la(pck), pair_lb(pck));
    if (cp.nq <= 0) return kOk;
    if (tn.block % 64 != 0 || tn.block <= 0 || tn.block > 256) {
        set_error(kEInval, "launch block %d", tn.block);
        return kEInval;
    }
    const long long qwaves = (cp.nq + 63) / 64;
    if (mode == kModeRaw || mode == kModeJK || mode == kModeEriOnly) {
        double* dst = (mode == kModeRaw) ? cp.out : cp.raw;
        const unsigned nb = gen_blocks(qwaves * g.n[0] * g.n[1], tn);
        switch (qcls) {
#define X(id, a, b, c, d)                                                                         \
    case id:                                                                                      \
        hipLaunchKernelGGL((k_eri_gen<a, b, c, d>), dim3(nb), dim3(tn.block), 0, s, cp, dst);     \
        break;
            SF_FOR_EACH_GENCLASS(X)
#undef X
        default:
            set_error(kEInval, "quartet class %d has no run-time-L kernel", qcls);
            return kEInval;
        }
        SF_HIP(hipGetLastError());
    }
    if (mode == kModeJK || mode == kModeDigestOnly) {
        hipLaunchKernelGGL(k_digest_gen, dim3(gen_blocks(qwaves * g.ntile, tn)), dim3(tn.block), 0,
                           s, cp, g, (const double*)cp.raw);
        SF_HIP(hipGetLastError());
    }
    return kOk;
}

int launch_schwarz_generic(int pc, const int* sa, const int* sb, const int* poff, const int* np,
                           int n, const PrimPair* prim, const double* shell_xyz, double* Q,
                           double* Qp, hipStream_t s)
{
    if (n <= 0) return kOk;
    const GenClass g = gen_class(pair_la(pc), pair_lb(pc), pair_la(pc), pair_lb(pc));
    hipLaunchKernelGGL(k_schwarz_gen, dim3((n + 63) / 64), dim3(64), 0, s, sa, sb, poff, np, n,
                       prim, shell_xyz, g, Q, Qp);
    SF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace sf
