// k_util.hip -- batched Boys / Rys kernels, the Schwarz quartet-list builder,
// the digestion plan and its deterministic (atomics-free) reduction, and the
// generic per-component quartet kernel behind the rys.eri_rys drop-in.
#include "sf_internal.hpp"

namespace sf {

// --------------------------------------------------------------- Boys -----
__global__ __launch_bounds__(256) void k_boys(long long n, int mmax, const double* __restrict__ T,
                                              double* __restrict__ F)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = T[i];
    if (!(t >= 0.0) || isinf(t)) {  // negative, NaN or infinite argument
        for (int m = 0; m <= mmax; ++m) F[i * (mmax + 1) + m] = (t == INFINITY) ? 0.0 : NAN;
        return;
    }
    double f[SF_BOYS_MTAB];
    boys_grid(mmax, t, f);
    for (int m = 0; m <= mmax; ++m) F[i * (mmax + 1) + m] = f[m];
}

int launch_boys(long long n, int mmax, const double* T, double* F, hipStream_t s)
{
    if (mmax < 0 || mmax > SF_BOYS_MTAB - 10) {
        set_error(kEInval, "mmax %d outside [0, %d]", mmax, SF_BOYS_MTAB - 10);
        return kEInval;
    }
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_boys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, mmax, T, F);
    SF_HIP(hipGetLastError());
    return kOk;
}

// ---------------------------------------------------------------- Rys -----
template <int N>
__global__ __launch_bounds__(256) void k_rys(long long n, const double* __restrict__ X,
                                             double* __restrict__ U, double* __restrict__ W)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[N], w[N];
    const double x = X[i];
    if (!(x >= 0.0)) {  // negative or NaN argument: no table lookup
#pragma unroll
        for (int r = 0; r < N; ++r) U[i * N + r] = W[i * N + r] = NAN;
        return;
    }
    rys_roots<N>(x, u, w);
#pragma unroll
    for (int r = 0; r < N; ++r) {
        U[i * N + r] = u[r];
        W[i * N + r] = w[r];
    }
}

int launch_rys(long long n, int nroots, const double* X, double* u, double* w, hipStream_t s)
{
    if (n <= 0) return kOk;
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    switch (nroots) {
    case 1: hipLaunchKernelGGL(k_rys<1>, g, b, 0, s, n, X, u, w); break;
    case 2: hipLaunchKernelGGL(k_rys<2>, g, b, 0, s, n, X, u, w); break;
    case 3: hipLaunchKernelGGL(k_rys<3>, g, b, 0, s, n, X, u, w); break;
    case 4: hipLaunchKernelGGL(k_rys<4>, g, b, 0, s, n, X, u, w); break;
    case 5: hipLaunchKernelGGL(k_rys<5>, g, b, 0, s, n, X, u, w); break;
    case 6: hipLaunchKernelGGL(k_rys<6>, g, b, 0, s, n, X, u, w); break;
    case 7: hipLaunchKernelGGL(k_rys<7>, g, b, 0, s, n, X, u, w); break;
    default: set_error(kEInval, "nroots %d outside [1, 7]", nroots); return kEInval;
    }
    SF_HIP(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------ quartet list ------
// For bra i (Q sorted descending within each molecule in both lists): the
// surviving kets are the prefix of its molecule's kets, ket0[i] + j with
// j < cut and Qbra[i] * Qket[ket0[i] + j] >= tau (the product is monotone in
// Qket), limited to ket0[i] + j <= i when bra and ket are the same pair
// class.  The FP64 part is the prefix that also meets tau_mixed.
__device__ __forceinline__ int count_prefix(double qi, const double* __restrict__ Qk, int n, double tau)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (qi * Qk[mid] >= tau) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_count_kets(const double* __restrict__ Qb, int nb, const double* __restrict__ Qk,
                             const int* __restrict__ ket0, const int* __restrict__ nket, int same,
                             double tau, double tau_mixed, int* __restrict__ n_all,
                             int* __restrict__ n_fp64)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double qi = Qb[i];
    const double* Qm = Qk + ket0[i];
    const int nk = nket[i];
    int n = count_prefix(qi, Qm, nk, tau);
    const int own = i - ket0[i];  // the bra's own position among its molecule's kets
    if (same && n > own + 1) n = own + 1;
    int n64 = n;
    if (tau_mixed > tau) {
        n64 = count_prefix(qi, Qm, nk, tau_mixed);
        if (n64 > n) n64 = n;
    }
    n_all[i] = n;
    n_fp64[i] = n64;
}

int launch_count_kets(const double* Qbra, int nbra, const double* Qket, const int* ket0,
                      const int* nket, int same, double tau, double tau_mixed, int* n_all,
                      int* n_fp64, hipStream_t s)
{
    if (nbra <= 0) return kOk;
    hipLaunchKernelGGL(k_count_kets, dim3((nbra + 255) / 256), dim3(256), 0, s, Qbra, nbra, Qket,
                       ket0, nket, same, tau, tau_mixed, n_all, n_fp64);
    SF_HIP(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------ digestion plan ----
__device__ __forceinline__ int find_base(const long long* __restrict__ base, int n, long long q)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Six plan entries per quartet: destination key of each digestion tile
// (molecule-local dense numbering, see KeyMap) and the code
// class | tile | quartet that k_plan_positions decodes.
__global__ __launch_bounds__(256) void k_plan_entries(int qcls, const PairListDev bra,
                                                      const PairListDev ket,
                                                      const int* __restrict__ ket0,
                                                      const long long* __restrict__ qall_base,
                                                      long long nq, const KeyMap km,
                                                      unsigned int* __restrict__ keys,
                                                      unsigned long long* __restrict__ vals)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int i = find_base(qall_base, bra.n, q);
    const int j = ket0[i] + (int)(q - qall_base[i]);
    const int m = km.sh_mol[bra.sa[i]];
    const unsigned ns = (unsigned)km.mol_ns[m], kb = km.mol_kbase[m];
    const unsigned a = km.sh_loc[bra.sa[i]], b = km.sh_loc[bra.sb[i]];
    const unsigned c = km.sh_loc[ket.sa[j]], d = km.sh_loc[ket.sb[j]];
    const unsigned n2 = ns * ns;
    const unsigned k[6] = {kb + a * ns + b,      kb + c * ns + d,      kb + n2 + a * ns + c,
                           kb + n2 + a * ns + d, kb + n2 + b * ns + c, kb + n2 + b * ns + d};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        keys[6 * q + t] = k[t];
        vals[6 * q + t] = ((unsigned long long)qcls << 58) | ((unsigned long long)t << 55) |
                          (unsigned long long)q;
    }
}

int launch_plan_entries(int qcls, const PairListDev& bra, const PairListDev& ket, const int* ket0,
                        const long long* qall_base, long long nq, const KeyMap& km,
                        unsigned int* keys, unsigned long long* vals, hipStream_t s)
{
    if (nq <= 0) return kOk;
    hipLaunchKernelGGL(k_plan_entries, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, qcls,
                       bra, ket, ket0, qall_base, nq, km, keys, vals);
    SF_HIP(hipGetLastError());
    return kOk;
}

__global__ __launch_bounds__(256) void k_segments(const unsigned int* __restrict__ keys, long long n,
                                                  long long* __restrict__ seg_begin,
                                                  long long* __restrict__ seg_end)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned k = keys[p];
    if (p == 0 || keys[p - 1] != k) seg_begin[k] = p;
    if (p == n - 1 || keys[p + 1] != k) seg_end[k] = p + 1;
}

int launch_segments(const unsigned int* keys, long long n, long long* seg_begin, long long* seg_end,
                    hipStream_t s)
{
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_segments, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys, n,
                       seg_begin, seg_end);
    SF_HIP(hipGetLastError());
    return kOk;
}

// For sorted plan position p (destination key k, entry code v = class | tile
// | quartet): slot of that tile in the destination-sorted tile buffer,
// key_base[k] + (p - seg_begin[k]) * NE(k), stored where the ERI kernel of
// that class will look it up.
__global__ __launch_bounds__(256) void k_plan_positions(const unsigned int* __restrict__ keys,
                                                        const unsigned long long* __restrict__ vals,
                                                        long long n,
                                                        const long long* __restrict__ seg_begin,
                                                        const unsigned long long* __restrict__ key_base,
                                                        const int* __restrict__ key_ne,
                                                        const ClassTab* __restrict__ tab)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned k = keys[p];
    const unsigned long long v = vals[p];
    const int cls = (int)(v >> 58);
    const int t = (int)((v >> 55) & 7);
    const long long q = (long long)(v & ((1ULL << 55) - 1));
    tab[cls].pos[t * tab[cls].qstride + q] =
        (unsigned)(key_base[k] + (unsigned long long)(p - seg_begin[k]) * key_ne[k]);
}

int launch_plan_positions(const unsigned int* keys, const unsigned long long* vals, long long n,
                          const long long* seg_begin, const unsigned long long* key_base,
                          const int* key_ne, const ClassTab* tab, hipStream_t s)
{
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_plan_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys,
                       vals, n, seg_begin, key_base, key_ne, tab);
    SF_HIP(hipGetLastError());
    return kOk;
}

// The transposed ket arrays of one sorted pair list (KetView layout): one
// thread per pair j writes field f of its primitive k at
// kt[(f * kmax + k) * n + j] (zero past its np[j] primitives) and its two
// centres at kc[e * n + j].
__global__ __launch_bounds__(256) void k_ket_transpose(const PairListDev pl, int kmax,
                                                       const PrimPair* __restrict__ prim,
                                                       const double* __restrict__ primq,
                                                       const double* __restrict__ prims,
                                                       const double* __restrict__ xyz,
                                                       double* __restrict__ kt,
                                                       double* __restrict__ kc)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= pl.n) return;
    const long long n = pl.n, fs = (long long)kmax * n;
    const int o = pl.poff[j], np = pl.np[j];
    for (int k = 0; k < kmax; ++k) {
        double f[kKetFields] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (k < np) {
            const PrimPair& pp = prim[o + k];
            f[0] = pp.p;
            f[1] = pp.Px;
            f[2] = pp.Py;
            f[3] = pp.Pz;
            f[4] = pp.ip;
            f[5] = pp.Kp;
            f[kKetQ] = primq[o + k];
            f[kKetS] = prims[o + k];
        }
#pragma unroll
        for (int e = 0; e < kKetFields; ++e) kt[e * fs + k * n + j] = f[e];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        kc[e * n + j] = xyz[3 * pl.sa[j] + e];
        kc[(3 + e) * n + j] = xyz[3 * pl.sb[j] + e];
    }
}

int launch_ket_transpose(const PairListDev& pl, int kmax, const PrimPair* prim, const double* primq,
                         const double* prims, const double* xyz, double* kt, double* kc,
                         hipStream_t s)
{
    if (pl.n <= 0) return kOk;
    hipLaunchKernelGGL(k_ket_transpose, dim3((unsigned)((pl.n + 255) / 256)), dim3(256), 0, s, pl,
                       kmax, prim, primq, prims, xyz, kt, kc);
    SF_HIP(hipGetLastError());
    return kOk;
}

// One-node classes: a code per quartet of the class (class-wide index qa),
// bit 0 = far (every primitive quartet, screened or not, has
// X = p q / (p + q) |P - Q|^2 >= kFarX), bit 1 = in the FP64 segment of a
// mixed build (ket index below n64 of its bra).  Depends on the geometry
// only, so a quartet takes the same path in every build of the handle.
__global__ __launch_bounds__(256) void k_far_codes(const ClassParams cp, long long nq,
                                                   const int* __restrict__ n64,
                                                   unsigned char* __restrict__ code)
{
    const long long qa = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (qa >= nq) return;
    const int i = find_bra(cp.qall_base, cp.bra.n, qa);
    const int jr = (int)(qa - cp.qall_base[i]);
    const int j = cp.ket0[i] + jr;
    const PrimPair* pb = cp.prim + cp.bra.poff[i];
    const PrimPair* pk = cp.prim + cp.ket.poff[j];
    const int nb = cp.bra.np[i], nk = cp.ket.np[j];
    // X = p q / (p + q) R^2 >= kFarX, tested without the division as
    // p q R^2 >= kFarX (p + q); stops at the first near primitive quartet
    bool far = true;
    for (int kk = 0; kk < nk && far; ++kk) {
        const double Qx = pk[kk].Px, Qy = pk[kk].Py, Qz = pk[kk].Pz, q = pk[kk].p;
        for (int kb = 0; kb < nb; ++kb) {
            const double dx = pb[kb].Px - Qx, dy = pb[kb].Py - Qy, dz = pb[kb].Pz - Qz;
            const double p = pb[kb].p;
            if (p * q * (dx * dx + dy * dy + dz * dz) < kFarX * (p + q)) {
                far = false;
                break;
            }
        }
    }
    code[qa] = (unsigned char)((far ? 1 : 0) | (jr < n64[i] ? 2 : 0));
}

int launch_far_codes(const ClassParams& cp, long long nq, const int* n64, unsigned char* code,
                     hipStream_t s)
{
    if (nq <= 0) return kOk;
    hipLaunchKernelGGL(k_far_codes, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, cp, nq, n64,
                       code);
    SF_HIP(hipGetLastError());
    return kOk;
}

// Destination reduction, pass 1: one wavefront per chunk (four per
// 256-thread block).  A chunk's tiles are contiguous (NE doubles each), so
// lanes read them coalesced; a fixed strided assignment of tiles to lanes
// and a fixed xor butterfly make the sum bitwise reproducible without
// atomics.  A chunk that is its destination's only one writes the result.
template <int N1, int N2>
__device__ __forceinline__ void store_block(const DestList& dl, int d, const ShellMap& map,
                                            const double* acc, double* X, double* Y)
{
    const int s1 = dl.s1[d], s2 = dl.s2[d];
    double* dst = (dl.xy[d] == 0 ? X : Y) + map.row[s1] + map.col[s2];
    const long long ld = map.ld[s1];
#pragma unroll
    for (int k = 0; k < N1 * N2; ++k) dst[(k / N2) * ld + k % N2] = acc[k];
}

template <int N1, int N2>
__global__ __launch_bounds__(256) void k_reduce_chunks(const DestList dl,
                                                       const double* __restrict__ tiles,
                                                       const ShellMap map, double* __restrict__ X,
                                                       double* __restrict__ Y)
{
    constexpr int NE = N1 * N2;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= dl.nch) return;
    const int lane = threadIdx.x & 63;
    const int cnt = dl.ch_n[w];
    const double* src = tiles + dl.ch_src[w];
    double acc[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) acc[k] = 0.0;
    // each lane adds tiles lane, lane + 64, lane + 128, ... in that order;
    // U tiles' loads are issued before their adds
    constexpr int U = reduce_unroll(NE);
    for (int e0 = lane; e0 < cnt; e0 += 64 * U) {
        double v[U][NE];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + 64 * u;
            const double* t = src + (long long)e * NE;
#pragma unroll
            for (int k = 0; k < NE; ++k) v[u][k] = (e < cnt) ? t[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (e0 + 64 * u < cnt)
#pragma unroll
                for (int k = 0; k < NE; ++k) acc[k] += v[u][k];
    }
#pragma unroll
    for (int k = 0; k < NE; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0) {
        const int slot = dl.ch_part[w];
        if (slot < 0) {
            store_block<N1, N2>(dl, dl.ch_dest[w], map, acc, X, Y);
        } else {
#pragma unroll
            for (int k = 0; k < NE; ++k) dl.part[(long long)slot * NE + k] = acc[k];
        }
    }
}

// Pass 2: one wavefront per destination with several chunks adds their
// partial sums (lanes strided over the chunks, fixed butterfly).
template <int N1, int N2>
__global__ __launch_bounds__(256) void k_reduce_partials(const DestList dl, const ShellMap map,
                                                         double* __restrict__ X,
                                                         double* __restrict__ Y)
{
    constexpr int NE = N1 * N2;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= dl.nm) return;
    const int lane = threadIdx.x & 63;
    const double* src = dl.part + (long long)dl.m_pbase[w] * NE;
    const int cnt = dl.m_nch[w];
    double acc[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) acc[k] = 0.0;
    for (int e = lane; e < cnt; e += 64) {
#pragma unroll
        for (int k = 0; k < NE; ++k) acc[k] += src[(long long)e * NE + k];
    }
#pragma unroll
    for (int k = 0; k < NE; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0) store_block<N1, N2>(dl, dl.m_dest[w], map, acc, X, Y);
}

int launch_reduce(int n1, int n2, const DestList& dl, const double* tiles, const ShellMap& map,
                  double* X, double* Y, hipStream_t s)
{
    if (dl.nch <= 0) return kOk;
    dim3 g1((dl.nch + 3) / 4), g2((dl.nm + 3) / 4), b(256);
#define R_(a, c)                                                                              \
    if (n1 == a && n2 == c) {                                                                 \
        hipLaunchKernelGGL((k_reduce_chunks<a, c>), g1, b, 0, s, dl, tiles, map, X, Y);       \
        SF_HIP(hipGetLastError());                                                            \
        if (dl.nm > 0) {                                                                      \
            hipLaunchKernelGGL((k_reduce_partials<a, c>), g2, b, 0, s, dl, map, X, Y);        \
            SF_HIP(hipGetLastError());                                                        \
        }                                                                                     \
        return kOk;                                                                           \
    }
    R_(1, 1) R_(1, 3) R_(1, 6) R_(3, 1) R_(3, 3) R_(3, 6) R_(6, 1) R_(6, 3) R_(6, 6)
    R_(1, 10) R_(3, 10) R_(6, 10) R_(10, 1) R_(10, 3) R_(10, 6) R_(10, 10)
#undef R_
    set_error(kEUnsupported, "tile %dx%d", n1, n2);
    return kEUnsupported;
}

// ------------------------------------------- gradient reductions ---------
// Per quartet q of a class (one thread each, flat over the class): the
// derivatives summed over the bra component slices, as two records of six
// doubles, ab[6q..] = (d/dA, d/dB) and kc[6q..] = (d/dC, d/dD = -(A+B+C)).
__global__ __launch_bounds__(256) void k_grad_slices(const double* __restrict__ gout, int ns,
                                                     long long qstride, double* __restrict__ ab,
                                                     double* __restrict__ kc)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= qstride) return;
    double gs[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < ns; ++s)
#pragma unroll
        for (int k = 0; k < 9; ++k) gs[k] += gout[(long long)(s * 9 + k) * qstride + q];
#pragma unroll
    for (int k = 0; k < 6; ++k) ab[6 * q + k] = gs[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        kc[6 * q + k] = gs[6 + k];
        kc[6 * q + 3 + k] = -(gs[k] + gs[3 + k] + gs[6 + k]);
    }
}

// Bra side: one wavefront per bra pair i of a class sums d/dA and d/dB over
// its kets (the contiguous quartets [base[i], base[i+1])), in a fixed order.
__global__ __launch_bounds__(256) void k_grad_reduce_bra(const double* __restrict__ ab,
                                                         const long long* __restrict__ base,
                                                         int nbra, double* __restrict__ part)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nbra) return;
    const int lane = threadIdx.x & 63;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long q = base[i] + lane; q < base[i + 1]; q += 64)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += ab[6 * q + k];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) part[6 * (long long)i + k] = acc[k];
}

// Ket side: one wavefront per ket pair j.  The bras holding ket j are the
// i in [i0, i1) with n_i = base[i+1] - base[i] > j (i0 = j when bra and ket
// lists coincide, else 0); n_i > j holds on a prefix of that range because
// the lists are sorted by Schwarz factor.  Sums d/dC and d/dD = -(A+B+C).
__global__ __launch_bounds__(256) void k_grad_reduce_ket(const double* __restrict__ kc,
                                                         const long long* __restrict__ base,
                                                         int nbra, int nket, int same,
                                                         double* __restrict__ part)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= nket) return;
    const int lane = threadIdx.x & 63;
    const int i0 = same ? j : 0;
    int lo = i0, hi = nbra;  // first i >= i0 with n_i <= j
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (base[mid + 1] - base[mid] > j) lo = mid + 1;
        else hi = mid;
    }
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = i0 + lane; i < lo; i += 64) {
        const double* v = kc + 6 * (base[i] + j);
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += v[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) part[6 * (long long)j + k] = acc[k];
}

// Per atom: fixed-order sum of the 3-vectors listed for it (CSR).
__global__ __launch_bounds__(64) void k_grad_atoms(const double* __restrict__ part,
                                                   const long long* __restrict__ off,
                                                   const long long* __restrict__ idx, int natom,
                                                   double* __restrict__ grad)
{
    const int a = blockIdx.x;
    if (a >= natom) return;
    const int lane = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long e = off[a] + lane; e < off[a + 1]; e += 64) {
        const double* v = part + idx[e];
        acc[0] += v[0];
        acc[1] += v[1];
        acc[2] += v[2];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) grad[3 * a + k] = acc[k];
}

int launch_grad_reduce_bra(const double* gout, int nslice, long long qstride, const long long* base,
                           int nbra, double* part, double* ab, double* kc, hipStream_t s)
{
    if (nbra <= 0 || qstride <= 0) return kOk;
    if ((qstride + 255) / 256 > 0x7fffffffLL) {
        set_error(kEUnsupported, "gradient reduction: %lld quartets", qstride);
        return kEUnsupported;
    }
    hipLaunchKernelGGL(k_grad_slices, dim3((unsigned)((qstride + 255) / 256)), dim3(256), 0, s, gout,
                       nslice, qstride, ab, kc);
    SF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_grad_reduce_bra, dim3((nbra + 3) / 4), dim3(256), 0, s,
                       (const double*)ab, base, nbra, part);
    SF_HIP(hipGetLastError());
    return kOk;
}

int launch_grad_reduce_ket(const double* kc, const long long* base, int nbra, int nket, int same,
                           double* part, hipStream_t s)
{
    if (nket <= 0) return kOk;
    hipLaunchKernelGGL(k_grad_reduce_ket, dim3((nket + 3) / 4), dim3(256), 0, s, kc, base, nbra,
                       nket, same, part);
    SF_HIP(hipGetLastError());
    return kOk;
}

int launch_grad_atoms(const double* part, const long long* csr_off, const long long* csr_idx,
                      int natom, double* grad, hipStream_t s)
{
    if (natom <= 0) return kOk;
    hipLaunchKernelGGL(k_grad_atoms, dim3(natom), dim3(64), 0, s, part, csr_off, csr_idx, natom,
                       grad);
    SF_HIP(hipGetLastError());
    return kOk;
}

// Column sums of the per-shell-pair one-electron gradient rows, plus `add`.
__global__ __launch_bounds__(64) void k_grad_rows(const double* __restrict__ rows, long long nrows,
                                                  int ncol, const double* __restrict__ add,
                                                  double* __restrict__ out)
{
    const int k = blockIdx.x;
    if (k >= ncol) return;
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (long long r = lane; r < nrows; r += 64) acc += rows[r * ncol + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) out[k] = acc + (add ? add[k] : 0.0);
}

int launch_grad_rows(const double* rows, long long nrows, int ncol, const double* add, double* out,
                     hipStream_t s)
{
    if (ncol <= 0) return kOk;
    hipLaunchKernelGGL(k_grad_rows, dim3(ncol), dim3(64), 0, s, rows, nrows, ncol, add, out);
    SF_HIP(hipGetLastError());
    return kOk;
}

// --------------------------------------------------- small dense kernels --
// Per-molecule element loops: blockIdx.y = molecule, x over its n^2 elements.
static int mol_grid(const MolLayout& ml, dim3* g)
{
    const long long per = (long long)ml.nmax * ml.nmax;
    if (ml.nmol <= 0 || ml.nmol > 65535 || per <= 0) {
        set_error(kEUnsupported, "layout of %d molecules (at most 65535)", ml.nmol);
        return kEUnsupported;
    }
    *g = dim3((unsigned)((per + 255) / 256), (unsigned)ml.nmol);
    return kOk;
}

__global__ void k_scale_density(const double* __restrict__ P, const double* __restrict__ f,
                                const MolLayout ml, double* __restrict__ Ps)
{
    const int m = blockIdx.y;
    const int n = ml.n[m];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const double* fm = f + ml.fbase[m];
    const long long idx = ml.off[m] + (long long)i * ml.ld[m] + j;
    Ps[idx] = fm[i] * fm[j] * P[idx];
}

int launch_scale_density(const double* P, const double* fnorm, const MolLayout& ml, double* Ps,
                         hipStream_t s)
{
    dim3 g;
    const int rc = mol_grid(ml, &g);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scale_density, g, dim3(256), 0, s, P, fnorm, ml, Ps);
    SF_HIP(hipGetLastError());
    return kOk;
}

// J = D (X + X^T) D, K = D (Y + Y^T) D per molecule block (D = per-component
// normalisation factors).
__global__ void k_finalize(const double* __restrict__ X, const double* __restrict__ Y,
                           const double* __restrict__ f, const MolLayout ml, double* __restrict__ J,
                           double* __restrict__ K)
{
    const int m = blockIdx.y;
    const int n = ml.n[m];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const double* fm = f + ml.fbase[m];
    const long long idx = ml.off[m] + (long long)i * ml.ld[m] + j;
    const long long tr = ml.off[m] + (long long)j * ml.ld[m] + i;
    const double sc = fm[i] * fm[j];
    J[idx] = sc * (X[idx] + X[tr]);
    K[idx] = sc * (Y[idx] + Y[tr]);
}

int launch_finalize(const double* X, const double* Y, const double* fnorm, const MolLayout& ml,
                    double* J, double* K, hipStream_t s)
{
    dim3 g;
    const int rc = mol_grid(ml, &g);
    if (rc) return rc;
    hipLaunchKernelGGL(k_finalize, g, dim3(256), 0, s, X, Y, fnorm, ml, J, K);
    SF_HIP(hipGetLastError());
    return kOk;
}

__global__ void k_fill(double* __restrict__ p, long long n, double v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill(double* p, long long n, double v, hipStream_t s)
{
    if (n <= 0) return kOk;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    SF_HIP(hipGetLastError());
    return kOk;
}

// -------------------------------------- component quartets (rys.eri_rys) --
// One thread per contracted Cartesian-component quartet, with the argument
// meaning of rys.eri_rys (src/hf/core/core.f90:967-1002): per-primitive
// normalisation factors and contraction coefficients, per-function
// (lx, ly, lz).  Total angular momentum L <= 13 with floor(L/2) + 1 <= 7
// nodes: the reference's 5 nodes (root5) up to L = 9, where they are exact;
// above, the exact rule instead of the reference's 5-node approximation.
template <int N>
__device__ double component_prim(const int* l, const double* xyz, const double* ex)
{
    const double a = ex[0], b = ex[1], c = ex[2], d = ex[3];
    const double p = a + b, q = c + d, pq = p + q, ipq = 1.0 / pq, rho = p * q * ipq;
    double P[3], Q[3], PQ2 = 0.0, Kab = 0.0, Kcd = 0.0;
    for (int k = 0; k < 3; ++k) {
        P[k] = (a * xyz[k] + b * xyz[3 + k]) / p;
        Q[k] = (c * xyz[6 + k] + d * xyz[9 + k]) / q;
        PQ2 += (P[k] - Q[k]) * (P[k] - Q[k]);
        Kab += (xyz[k] - xyz[3 + k]) * (xyz[k] - xyz[3 + k]);
        Kcd += (xyz[6 + k] - xyz[9 + k]) * (xyz[6 + k] - xyz[9 + k]);
    }
    const double pref = 34.986836655249725 * exp(-a * b / p * Kab - c * d / q * Kcd) / (p * q * sqrt(pq));
    double u[N], w[N];
    rys_roots<N>(rho * PQ2, u, w);
    double res = 0.0;
    for (int r = 0; r < N; ++r) {
        const double t2 = u[r];
        const double b00 = 0.5 * ipq * t2, b10 = 0.5 / p * (1.0 - q * ipq * t2),
                     b01 = 0.5 / q * (1.0 - p * ipq * t2);
        double prod = pref * w[r];
        for (int k = 0; k < 3; ++k) {
            const int li = l[k], lj = l[3 + k], lk = l[6 + k], ll = l[9 + k];
            const int nn = li + lj, mm = lk + ll;
            double G[14][14];  // total L <= 13 (checked by the caller)
            const double c00 = (P[k] - xyz[k]) - q * ipq * (P[k] - Q[k]) * t2;
            const double c00p = (Q[k] - xyz[6 + k]) + p * ipq * (P[k] - Q[k]) * t2;
            G[0][0] = 1.0;
            if (nn > 0) G[1][0] = c00;
            for (int n = 1; n < nn; ++n) G[n + 1][0] = n * b10 * G[n - 1][0] + c00 * G[n][0];
            for (int m = 0; m < mm; ++m)
                for (int n = 0; n <= nn; ++n) {
                    double g = c00p * G[n][m];
                    if (m > 0) g += m * b01 * G[n][m - 1];
                    if (n > 0) g += n * b00 * G[n - 1][m];
                    G[n][m + 1] = g;
                }
            const double xij = xyz[k] - xyz[3 + k], xkl = xyz[6 + k] - xyz[9 + k];
            double s = 0.0;
            for (int m = 0; m <= ll; ++m) {
                double t = 0.0;
                for (int n = 0; n <= lj; ++n)
                    t += binom_c(lj, n) * pow(xij, (double)(lj - n)) * G[n + li][m + lk];
                s += binom_c(ll, m) * pow(xkl, (double)(ll - m)) * t;
            }
            prod *= s;
        }
        res += prod;
    }
    return res;
}

__global__ __launch_bounds__(128) void k_eri_components(
    const int* __restrict__ npgf, const int* __restrict__ poff, const double* __restrict__ coefs,
    const double* __restrict__ expns, const double* __restrict__ norms,
    const int* __restrict__ shell, const double* __restrict__ center, long long nq,
    const int* __restrict__ quartets, double* __restrict__ out)
{
    const long long qi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    int f[4], l[12], L = 0;
    double xyz[12];
    for (int k = 0; k < 4; ++k) {
        f[k] = quartets[4 * qi + k];
        for (int e = 0; e < 3; ++e) {
            l[3 * k + e] = shell[3 * f[k] + e];
            xyz[3 * k + e] = center[3 * f[k] + e];
            L += l[3 * k + e];
        }
    }
    const int nr = L / 2 + 1;
    double res = 0.0;
    for (int i1 = 0; i1 < npgf[f[0]]; ++i1)
        for (int i2 = 0; i2 < npgf[f[1]]; ++i2)
            for (int i3 = 0; i3 < npgf[f[2]]; ++i3)
                for (int i4 = 0; i4 < npgf[f[3]]; ++i4) {
                    const int p1 = poff[f[0]] + i1, p2 = poff[f[1]] + i2, p3 = poff[f[2]] + i3,
                              p4 = poff[f[3]] + i4;
                    const double ex[4] = {expns[p1], expns[p2], expns[p3], expns[p4]};
                    double v;
                    switch (nr) {
                    case 1: v = component_prim<1>(l, xyz, ex); break;
                    case 2: v = component_prim<2>(l, xyz, ex); break;
                    case 3: v = component_prim<3>(l, xyz, ex); break;
                    case 4: v = component_prim<4>(l, xyz, ex); break;
                    case 5: v = component_prim<5>(l, xyz, ex); break;
                    case 6: v = component_prim<6>(l, xyz, ex); break;
                    default: v = component_prim<7>(l, xyz, ex); break;
                    }
                    res += norms[p1] * norms[p2] * norms[p3] * norms[p4] * coefs[p1] * coefs[p2] *
                           coefs[p3] * coefs[p4] * v;
                }
    out[qi] = res;
}

int launch_eri_components(int nf, const int* npgf, const int* poff, const double* coefs,
                          const double* expns, const double* norms, const int* shell,
                          const double* center, long long nq, const int* quartets, double* out,
                          hipStream_t s)
{
    (void)nf;
    if (nq <= 0) return kOk;
    hipLaunchKernelGGL(k_eri_components, dim3((unsigned)((nq + 127) / 128)), dim3(128), 0, s, npgf,
                       poff, coefs, expns, norms, shell, center, nq, quartets, out);
    SF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace sf
