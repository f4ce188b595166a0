// k_onee.hip -- one-electron integrals S, T, V on the GPU (McMurchie-Davidson).
//
// Reference: mmd2.overlap / kinetic / potential (src/hf/mmd2.py:14-84) over
// the Fortran mmd module (core.f90:30-97), assembled per function pair in
// Python loops at hf.py:266-274.  Here one thread computes one shell pair
// (all components, all primitive pairs, all nuclei) and writes its block
// and the transposed block: every matrix element has exactly one writer.
//   S_ab = K (pi/p)^{3/2} E^x_0 E^y_0 E^z_0,
//   T_ab = sum_k [-2 b^2 S^k_{a,b+2} + b (2 b_k + 1) S^k_{ab} - b_k(b_k-1)/2 S^k_{a,b-2}] prod_{k'!=k} S^k'
//          (mmd2.kinetic_3d, mmd2.py:31-42),
//   V_ab = sum_C -Z_C (2 pi / p) K sum_{tuv} E^x_t E^y_u E^z_v R_{tuv}(p, P - C)
//          (mmd2.potential_3d, mmd2.py:53-76; the -Z_C factor of hf.py:85-88).
#include "sf_internal.hpp"

namespace sf {

template <int LA, int LB>
__global__ __launch_bounds__(128) void k_one_electron(const int* __restrict__ sa, const int* __restrict__ sb,
                                                      const int* __restrict__ poff, const int* __restrict__ np,
                                                      int n, const PrimPair* __restrict__ prim,
                                                      const double* __restrict__ xyz,
                                                      const int* __restrict__ foff,
                                                      const double* __restrict__ fnorm,
                                                      const ShellMap map,
                                                      const int* __restrict__ sh_mol,
                                                      const int* __restrict__ mol_nuc,
                                                      const double* __restrict__ Zc,
                                                      const double* __restrict__ Cxyz,
                                                      double* __restrict__ S, double* __restrict__ T,
                                                      double* __restrict__ V)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    // nuclei of the pair's own molecule only
    const int m = sh_mol[a];
    const int c0n = mol_nuc[m], nn_c = mol_nuc[m + 1] - c0n;
    double s[NA * NB], t[NA * NB], v[NA * NB];
    one_electron_shell_pair<LA, LB>(prim + poff[i], np[i], A, B, nn_c, Zc + c0n, Cxyz + 3 * c0n, s, t, v);
    const int r0 = foff[a], c0 = foff[b];
    const long long ld = map.ld[a];
#pragma unroll
    for (int fa = 0; fa < NA; ++fa)
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
            const double nn = fnorm[r0 + fa] * fnorm[c0 + fb];
            const long long ij = map.row[a] + fa * ld + map.col[b] + fb;
            const long long ji = map.row[b] + fb * ld + map.col[a] + fa;
            S[ij] = S[ji] = nn * s[fa * NB + fb];
            T[ij] = T[ji] = nn * t[fa * NB + fb];
            V[ij] = V[ji] = nn * v[fa * NB + fb];
        }
}

int launch_one_electron(int pc, const int* sa, const int* sb, const int* poff, const int* np, int n,
                        const PrimPair* prim, const double* xyz, const int* foff,
                        const double* fnorm, const ShellMap& map, const int* sh_mol,
                        const int* mol_nuc, const double* Z, const double* Cxyz, double* S,
                        double* T, double* V, hipStream_t s)
{
    if (n <= 0) return kOk;
    dim3 g((n + 127) / 128), b(128);
#define O_(id, la, lb)                                                                          \
    case id:                                                                                    \
        hipLaunchKernelGGL((k_one_electron<la, lb>), g, b, 0, s, sa, sb, poff, np, n, prim, xyz, \
                           foff, fnorm, map, sh_mol, mol_nuc, Z, Cxyz, S, T, V);               \
        break;
    switch (pc) {
        O_(0, 0, 0) O_(1, 1, 0) O_(2, 1, 1) O_(3, 2, 0) O_(4, 2, 1) O_(5, 2, 2)
        O_(6, 3, 0) O_(7, 3, 1) O_(8, 3, 2) O_(9, 3, 3)
    default: set_error(kEUnsupported, "pair class %d", pc); return kEUnsupported;
    }
#undef O_
    SF_HIP(hipGetLastError());
    return kOk;
}

// Gradient of the one-electron energy (s, p, d shells), in two passes over a pair class:
//  k_grad_onee_v:  one thread per (shell pair, nucleus c) computes the
//                  attraction derivatives for c: writes d/dC into the pair's
//                  row at atom c and d/dA, d/dB into part[(i*natom + c)*6];
//  k_grad_onee_st: one thread per shell pair adds the overlap/kinetic
//                  derivatives and the per-nucleus d/dA, d/dB in nucleus
//                  order into the row at the pair's atoms.
// Rows are summed by k_grad_rows in a fixed order (one_electron_grad_pair
// is the same computation on one thread).
template <int LA, int LB>
__device__ __forceinline__ void load_pair_blocks(const ShellMap& map, int a, int b,
                                                 const double* __restrict__ Ps,
                                                 const double* __restrict__ Ws, double* Pb,
                                                 double* Wb)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    const long long ld = map.ld[a];
#pragma unroll
    for (int fa = 0; fa < NA; ++fa)
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
            const long long ij = map.row[a] + fa * ld + map.col[b] + fb;
            Pb[fa * NB + fb] = Ps[ij];
            if (Wb) Wb[fa * NB + fb] = Ws[ij];
        }
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_grad_onee_v(const int* __restrict__ sa, const int* __restrict__ sb,
                                                    const int* __restrict__ poff, const int* __restrict__ np,
                                                    int n, const PrimPair* __restrict__ prim,
                                                    const double* __restrict__ xyz, const ShellMap map,
                                                    int natom, const double* __restrict__ Zc,
                                                    const double* __restrict__ Cxyz,
                                                    const double* __restrict__ Ps,
                                                    double* __restrict__ rows, double* __restrict__ part)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * natom) return;
    const int i = (int)(t / natom), c = (int)(t % natom);
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    double Pb[NA * NB];
    load_pair_blocks<LA, LB>(map, a, b, Ps, nullptr, Pb, nullptr);
    double gA[3], gB[3], gC[3];
    one_electron_grad_v<LA, LB>(prim + poff[i], np[i], A, B, c, Zc, Cxyz, Pb, (a == b) ? 1.0 : 2.0,
                                gA, gB, gC);
    double* row = rows + (long long)i * 3 * natom;
    double* pp = part + ((long long)i * natom + c) * 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        row[3 * c + d] = gC[d];
        pp[d] = gA[d];
        pp[3 + d] = gB[d];
    }
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_grad_onee_st(const int* __restrict__ sa, const int* __restrict__ sb,
                                                     const int* __restrict__ poff, const int* __restrict__ np,
                                                     int n, const PrimPair* __restrict__ prim,
                                                     const double* __restrict__ xyz, const ShellMap map,
                                                     const int* __restrict__ shell_atom, int natom,
                                                     const double* __restrict__ Ps,
                                                     const double* __restrict__ Ws,
                                                     double* __restrict__ rows,
                                                     const double* __restrict__ part)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    double Pb[NA * NB], Wb[NA * NB];
    load_pair_blocks<LA, LB>(map, a, b, Ps, Ws, Pb, Wb);
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    one_electron_grad_st<LA, LB>(prim + poff[i], np[i], A, B, Pb, Wb, (a == b) ? 1.0 : 2.0, gA, gB);
    const double* pp = part + (long long)i * natom * 6;
    for (int c = 0; c < natom; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gA[d] += pp[6 * c + d];
            gB[d] += pp[6 * c + 3 + d];
        }
    double* row = rows + (long long)i * 3 * natom;
    const int atomA = shell_atom[a], atomB = shell_atom[b];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        row[3 * atomA + d] += gA[d];
        row[3 * atomB + d] += gB[d];
    }
}

int launch_grad_one_electron(int pc, const int* sa, const int* sb, const int* poff, const int* np,
                             int n, const PrimPair* prim, const double* xyz, const ShellMap& map,
                             const int* shell_atom, int natom, const double* Z, const double* Cxyz,
                             const double* Ps, const double* Ws, double* rows, double* part,
                             hipStream_t s)
{
    if (n <= 0 || natom <= 0) return kOk;
    const long long nt = (long long)n * natom;
    if ((nt + 63) / 64 > 0x7fffffffLL) {
        set_error(kEUnsupported, "one-electron gradient: %lld (pair, nucleus) items", nt);
        return kEUnsupported;
    }
    dim3 gv((unsigned)((nt + 63) / 64)), gs((n + 63) / 64), b(64);
#define O_(id, la, lb)                                                                            \
    case id:                                                                                      \
        hipLaunchKernelGGL((k_grad_onee_v<la, lb>), gv, b, 0, s, sa, sb, poff, np, n, prim, xyz,   \
                           map, natom, Z, Cxyz, Ps, rows, part);                                  \
        SF_HIP(hipGetLastError());                                                                \
        hipLaunchKernelGGL((k_grad_onee_st<la, lb>), gs, b, 0, s, sa, sb, poff, np, n, prim, xyz,  \
                           map, shell_atom, natom, Ps, Ws, rows, (const double*)part);            \
        break;
    switch (pc) {
        O_(0, 0, 0) O_(1, 1, 0) O_(2, 1, 1) O_(3, 2, 0) O_(4, 2, 1) O_(5, 2, 2)
    default: set_error(kEUnsupported, "pair class %d", pc); return kEUnsupported;
    }
#undef O_
    SF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace sf
