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
    default: set_error(kEUnsupported, "pair class %d", pc); return kEUnsupported;
    }
#undef O_
    SF_HIP(hipGetLastError());
    return kOk;
}

// Gradient of the one-electron energy: one thread per shell pair writes its
// row of 3*natom derivatives (one_electron_grad_pair); rows are summed by
// k_grad_rows in a fixed order.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_grad_onee(const int* __restrict__ sa, const int* __restrict__ sb,
                                                  const int* __restrict__ poff, const int* __restrict__ np,
                                                  int n, const PrimPair* __restrict__ prim,
                                                  const double* __restrict__ xyz,
                                                  const ShellMap map,
                                                  const int* __restrict__ shell_atom, int natom,
                                                  const double* __restrict__ Zc,
                                                  const double* __restrict__ Cxyz,
                                                  const double* __restrict__ Ps,
                                                  const double* __restrict__ Ws,
                                                  double* __restrict__ rows)
{
    constexpr int NA = cart_n(LA), NB = cart_n(LB);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = sa[i], b = sb[i];
    const double A[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const double B[3] = {xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]};
    double Pb[NA * NB], Wb[NA * NB];
    const long long ld = map.ld[a];
#pragma unroll
    for (int fa = 0; fa < NA; ++fa)
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
            const long long ij = map.row[a] + fa * ld + map.col[b] + fb;
            Pb[fa * NB + fb] = Ps[ij];
            Wb[fa * NB + fb] = Ws[ij];
        }
    one_electron_grad_pair<LA, LB>(prim + poff[i], np[i], A, B, shell_atom[a], shell_atom[b], natom,
                                   Zc, Cxyz, Pb, Wb, (a == b) ? 1.0 : 2.0,
                                   rows + (long long)i * 3 * natom);
}

int launch_grad_one_electron(int pc, const int* sa, const int* sb, const int* poff, const int* np,
                             int n, const PrimPair* prim, const double* xyz, const ShellMap& map,
                             const int* shell_atom, int natom, const double* Z, const double* Cxyz,
                             const double* Ps, const double* Ws, double* rows, hipStream_t s)
{
    if (n <= 0) return kOk;
    dim3 g((n + 63) / 64), b(64);
#define O_(id, la, lb)                                                                            \
    case id:                                                                                      \
        hipLaunchKernelGGL((k_grad_onee<la, lb>), g, b, 0, s, sa, sb, poff, np, n, prim, xyz, map,  \
                           shell_atom, natom, Z, Cxyz, Ps, Ws, rows);                             \
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
