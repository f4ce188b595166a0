// sf_internal.hpp -- launch parameter blocks and launcher declarations shared
// by the kernel translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "eri_gen.hpp"
#include "mmd.hpp"

namespace sf {

// Status codes returned by every C-ABI entry point.
enum Status : int {
    kOk = 0,
    kEInval = 1,      // bad argument
    kEHip = 2,        // HIP runtime error
    kENoMem = 3,      // allocation failed
    kEState = 4,      // call made in the wrong order (e.g. no screening yet)
    kEUnsupported = 5 // angular momentum or size beyond what is built
};

void set_error(int code, const char* fmt, ...);

#define SF_HIP(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            ::sf::set_error(::sf::kEHip, "%s:%d %s: %s", __FILE__, __LINE__, #x,         \
                            hipGetErrorString(e_));                                      \
            return ::sf::kEHip;                                                          \
        }                                                                                \
    } while (0)

constexpr int kMaxL = 3;            // s, p, d, f shells
constexpr int kNumPairClass = 10;   // (00) (10) (11) (20) (21) (22) (30) (31) (32) (33)
constexpr int kNumQuartetClass = 55;
// Quartet classes 0..20 hold s, p, d shells only, classes 21..54 an f
// shell; pair classes 0..5 s, p, d, 6..9 an f shell.  Every s/p/d class and
// the f classes of up to 540 integrals per quartet have compiled kernels
// (k_eri.hip); the larger f classes run the run-time-L kernels of
// k_eri_gen.hip (class_generic).
constexpr int kNumClassSPD = 21;
constexpr int kNumPairSPD = 6;
// Digestion tile shapes N1 x N2 with N1, N2 in {1, 3, 6, 10} (s, p, d, f
// components): tile group g = 4 L1 + L2.
constexpr int kNumTileGroup = 16;
// sf_jk_profile: milliseconds per launch group, [2c] / [2c + 1] the FP64 /
// FP32 segment of quartet class c, then density scaling and the reduction.
constexpr int kSlotScale = 2 * kNumQuartetClass;
constexpr int kSlotReduce = kSlotScale + 1;
constexpr int kNumSlots = kSlotScale + 2;

constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }  // la >= lb
constexpr int pair_la(int pc) { return pc == 0 ? 0 : pc < 3 ? 1 : pc < 6 ? 2 : 3; }
constexpr int pair_lb(int pc) { return pc - pair_la(pc) * (pair_la(pc) + 1) / 2; }
constexpr int quartet_class(int pcb, int pck) { return pcb * (pcb + 1) / 2 + pck; }  // pcb >= pck
// bra pair class of quartet class c
constexpr int pair_of_class(int c)
{
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= c) ++p;
    return p;
}
// total angular momentum la + lb + lc + ld of quartet class c
constexpr int class_ltot(int c)
{
    const int pb = pair_of_class(c), pk = c - pb * (pb + 1) / 2;
    return pair_la(pb) + pair_lb(pb) + pair_la(pk) + pair_lb(pk);
}
static_assert(class_ltot(0) == 0 && class_ltot(1) == 1 && class_ltot(2) == 2, "class order");
static_assert(quartet_class(pair_class(3, 0), 0) == kNumClassSPD, "f classes follow s/p/d");
static_assert(quartet_class(pair_class(3, 3), pair_class(3, 3)) == kNumQuartetClass - 1, "55 classes");

// One sorted pair class on the device.
struct PairListDev {
    const int* sa;    // first shell (L(sa) >= L(sb))
    const int* sb;
    const int* poff;  // first primitive pair
    const int* np;    // number of primitive pairs
    int n;
};

// Where each shell's functions live in the N x N matrices (P, J, K, S, T, V)
// of its molecule: component i of shell a and component j of shell b (same
// molecule) are at row[a] + i * ld[a] + col[b] + j.  With one molecule,
// row = foff * nbf, col = foff, ld = nbf; a batch of molecules puts each
// molecule's block wherever the caller's layout says (see sf_basis_create_batch).
struct ShellMap {
    const long long* row;
    const int* col;
    const int* ld;
};

// Launch shape of a kernel: `block` threads per block; `grid` = 0 launches
// one thread (or wavefront) per work item, grid > 0 launches at most that
// many blocks, which then stride over the work (persistent blocks).
struct Tune {
    int block = 256;
    int grid = 0;
    // FP32 segment of a mixed build only: 1 = run its quartets with FP64
    // arithmetic (same primitive threshold), where that is the faster one;
    // the class then runs both segments as one launch (ClassParams::n64)
    int arith64 = 0;
};

// One launch of a quartet-class kernel over a precision segment.
struct ClassParams {
    PairListDev bra, ket;
    const PrimPair* prim;
    const double* shell_xyz;   // 3 per shell
    ShellMap map;              // matrix positions of each shell's functions
    const long long* seg_base; // nbra+1: flat segment index of each bra's first quartet
    const int* seg_k0;         // nbra: first ket of the segment, relative to ket0
    const int* ket0;           // nbra: first ket of the bra's molecule in the ket list
    const long long* qall_base;// nbra+1: class-wide quartet index of ket 0 of each bra
    long long nq;              // quartets in this segment
    long long qstride;         // quartets in the class (plane stride of out)
    int same;                  // bra and ket pair classes coincide
    const double* P;           // density in the shell-normalised basis
    const double* primq;       // Schwarz factor of each primitive pair (same index as prim)
    double tau_prim;           // skip primitive quartets with Qp_b Qp_k < tau_prim (per segment)
    const int* n64;            // one launch for both mixed segments: kets at index >= n64[i]
    double tau_prim_lo;        //   of bra i use tau_prim_lo (nullptr: tau_prim for all)
    double* out;               // kModeRaw: SoA planes out[k * qstride + q]
    double* raw;               // split classes: raw integral planes (NINT * qstride)
    const unsigned* pos;       // kModeJK: pos[t * qstride + q] = slot of tile t in `tiles`
    double* tiles;             // kModeJK: destination-sorted digestion tiles
    // one-node classes ((ss|ss), (ps|ss)), kModeJK: with qlist, the launch
    // covers the class-wide quartet indices qlist[0..nq) instead of the
    // segment; with far, every primitive quartet of those quartets has
    // X >= kFarX and the point-charge form runs (prims: S = pi^{3/2} K p^{-3/2}
    // of each primitive pair, same index as prim)
    const unsigned* qlist = nullptr;
    int far = 0;
    const double* prims = nullptr;
    // the ket pair list's primitive pairs transposed (KetView: field stride
    // kt_fs = kmax * kt_n, primitive stride kt_n) and its pairs' centres
    // (kc[k * kt_n + j], Cx Cy Cz Dx Dy Dz)
    const double* kt = nullptr;
    long long kt_fs = 0, kt_n = 0;
    const double* kc = nullptr;
};

// Blocks of a launch are dispatched to the 8 XCDs round robin (block p on
// XCD p % 8), each XCD with its own L2.  xcd_block() renumbers them so that
// runs of kXcdRun consecutive logical blocks stay on one XCD: neighbouring
// quartets (mostly one bra's), whose digestion tiles share cache lines, are
// written through one L2, where their partial-line stores merge, while the
// runs still interleave over the XCDs (the cost per quartet drifts along a
// launch).  The blocks past the last whole round of runs are split into one
// contiguous range per XCD.  A bijection for any grid size.
constexpr unsigned kXcd = 8, kXcdRun = 64;
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned p = blockIdx.x, nb = gridDim.x, x = p % kXcd, k = p / kXcd;
    const unsigned round = kXcd * kXcdRun, whole = nb / round * round;
    if (p < whole) return k / kXcdRun * round + x * kXcdRun + k % kXcdRun;
    const unsigned nt = nb - whole, q = nt / kXcd, r = nt % kXcd;
    return whole + x * q + (x < r ? x : r) + (p - whole) / kXcd;
}

// Flat segment index q -> bra i, ket j (index in the ket list) and the
// class-wide quartet index qa.
struct QuartetIdx {
    int i, j;
    long long qa;
};

__device__ __forceinline__ int find_bra(const long long* __restrict__ base, int n, long long q)
{
    int lo = 0, hi = n;  // base[lo] <= q < base[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The bra of quartet q: one binary search per wavefront for its first
// lane's quartet qw (wave-uniform, so scalar loads), then each lane walks
// forward over the few bras the wavefront spans.  Requires qw <= q < nq.
// The same for a launch over a list of class-wide quartet indices (sorted
// ascending): the wavefront searches for the bra of its first lane's entry.
__device__ __forceinline__ QuartetIdx locate_list(const ClassParams& cp, long long t, long long tw)
{
    const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)(tw & 0xffffffffLL));
    const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(tw >> 32));
    const long long tu = ((long long)hi32 << 32) | lo32;
    int i = __builtin_amdgcn_readfirstlane(find_bra(cp.qall_base, cp.bra.n, (long long)cp.qlist[tu]));
    const long long qa = cp.qlist[t];
    while (cp.qall_base[i + 1] <= qa) ++i;
    QuartetIdx r;
    r.i = i;
    r.j = cp.ket0[i] + (int)(qa - cp.qall_base[i]);
    r.qa = qa;
    return r;
}

__device__ __forceinline__ QuartetIdx locate(const ClassParams& cp, long long q, long long qw)
{
    const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)(qw & 0xffffffffLL));
    const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(qw >> 32));
    const long long qu = ((long long)hi32 << 32) | lo32;
    int i = __builtin_amdgcn_readfirstlane(find_bra(cp.seg_base, cp.bra.n, qu));
    while (cp.seg_base[i + 1] <= q) ++i;
    QuartetIdx r;
    r.i = i;
    const int jr = cp.seg_k0[i] + (int)(q - cp.seg_base[i]);
    r.j = cp.ket0[i] + jr;
    r.qa = cp.qall_base[i] + jr;
    return r;
}


// kModeJK: integrals and digestion tiles; kModeRaw: raw integral planes
// into ClassParams::out (dense export).  Split classes only:
// kModeEriOnly writes the raw planes into ClassParams::raw without
// digesting them, kModeDigestOnly digests ClassParams::raw (so the two
// segments of a mixed build can share one digestion launch).
// kModeJKFar (kernel instantiation only): kModeJK for the far quartets of
// a one-node class (ClassParams::far), with the point-charge integrals.
enum EriMode : int { kModeJK = 0, kModeRaw = 1, kModeEriOnly = 2, kModeDigestOnly = 3,
                     kModeJKFar = 4 };

// Where the plan-position kernel writes the slots of one quartet class.
struct ClassTab {
    unsigned* pos;       // pos[t * qstride + q]
    long long qstride;
};

// Per-molecule layout of the N x N matrices: molecule m has n[m] functions
// (global function indices fbase[m] ...), its block starts at off[m] with
// leading dimension ld[m] >= n[m]; nmax = max n[m].
struct MolLayout {
    const long long* off;
    const int* ld;
    const int* n;
    const int* fbase;
    int nmol;
    int nmax;
};

// Digestion destinations of one tile shape (N1 x N2 blocks): the two
// shells of each destination block and which matrix it belongs to (0 = X
// for J, 1 = Y for K).  A destination's tiles (a contiguous run in the
// destination-sorted tile buffer) are summed in chunks of at most
// reduce_chunk_tiles(NE) tiles of NE = N1 N2 doubles, one wavefront per
// chunk; a destination with several chunks gets partial sums (slots
// ch_part) that a second pass adds in chunk order.  Both passes have a
// fixed order: the result is reproducible bit for bit.  Small tiles take
// long chunks (more loads in flight per wavefront), large ones short
// chunks (more wavefronts, fewer registers).
constexpr int reduce_chunk_tiles(int ne) { return ne == 1 ? 1024 : ne <= 3 ? 512 : 256; }
// tiles whose loads a lane issues before adding them
constexpr int reduce_unroll(int ne) { return ne <= 3 ? 4 : 1; }
struct DestList {
    const int* s1;
    const int* s2;
    const int* xy;
    const unsigned long long* ch_src;  // per chunk: first double of its first tile
    const int* ch_n;                   // tiles in the chunk
    const int* ch_dest;                // destination of the chunk
    const int* ch_part;                // partial slot, -1 = the chunk is its whole destination
    int nch;
    const int* m_dest;                 // destinations with several chunks
    const int* m_pbase;                //   their first partial slot
    const int* m_nch;                  //   and chunk count
    int nm;
    double* part;                      // partial sums, N1*N2 doubles per slot
};

// Shell -> molecule bookkeeping for plan keys: molecule of each shell, its
// index inside the molecule, shells per molecule, first plan key of each
// molecule (keys of molecule m: kbase[m] + xy * ns^2 + s1 * ns + s2).
struct KeyMap {
    const int* sh_mol;
    const int* sh_loc;
    const int* mol_ns;
    const unsigned* mol_kbase;
};

// Launchers (k_eri_gen.hip): the f-shell quartet classes without compiled
// kernels (class_generic) in every mode (integrals always in FP64), and the
// Schwarz factors of the f pair classes.
int launch_eri_generic(int qcls, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s);
int launch_schwarz_generic(int pc, const int* sa, const int* sb, const int* poff, const int* np,
                           int n, const PrimPair* prim, const double* shell_xyz, double* Q,
                           double* Qp, hipStream_t s);

// Launchers (k_eri.hip).
int launch_eri_class(int qcls, int fp32, int mode, const ClassParams& cp, const Tune& tn,
                     hipStream_t s);
// McMurchie-Davidson engine (MFMA E.R.E), FP64: raw planes (kModeRaw into
// cp.out) or raw planes in cp.raw plus digestion (kModeJK).
int launch_eri_mmd(int qcls, int mode, const ClassParams& cp, const Tune& tn, hipStream_t s);
int launch_schwarz(int pc, const int* sa, const int* sb, const int* poff, const int* np, int n,
                   const PrimPair* prim, const double* shell_xyz, double* Q, double* Qp,
                   hipStream_t s);
int class_ntile(int qcls);
// 1 for the f-shell classes that run the run-time-L kernels (k_eri_gen.hip)
int class_generic(int qcls);
int class_nint(int qcls);
int class_split(int qcls);
int class_nslice(int qcls);
int launch_grad_class(int qcls, int fp32, const ClassParams& cp, double* gout, hipStream_t s);
// per-block counts of the primitive quartets the Rys kernels evaluate in a segment
int launch_count_prims(const ClassParams& cp, int nblock, unsigned long long* part, hipStream_t s);
int launch_grad_reduce_bra(const double* gout, int nslice, long long qstride,
                           const long long* base, int nbra, double* part, double* ab,
                           double* kc, hipStream_t s);
int launch_grad_reduce_ket(const double* kc, const long long* base, int nbra, int nket, int same,
                           double* part, hipStream_t s);
int launch_grad_atoms(const double* part, const long long* csr_off, const long long* csr_idx,
                      int natom, double* grad, hipStream_t s);
int launch_grad_one_electron(int pc, const int* sa, const int* sb, const int* poff, const int* np,
                             int n, const PrimPair* prim, const double* xyz, const ShellMap& map,
                             const int* shell_atom, int natom, const double* Z, const double* Cxyz,
                             const double* Ps, const double* Ws, double* rows, double* part,
                             hipStream_t s);
int launch_grad_rows(const double* rows, long long nrows, int ncol, const double* add, double* out,
                     hipStream_t s);
void class_tile_offsets(int qcls, int off[6]);

// Launchers (k_util.hip).
int launch_boys(long long n, int mmax, const double* T, double* F, hipStream_t s);
int launch_rys(long long n, int nroots, const double* X, double* u, double* w, hipStream_t s);
int launch_count_kets(const double* Qbra, int nbra, const double* Qket, const int* ket0,
                      const int* nket, int same, double tau, double tau_mixed, int* n_all,
                      int* n_fp64, hipStream_t s);
int launch_plan_entries(int qcls, const PairListDev& bra, const PairListDev& ket, const int* ket0,
                        const long long* qall_base, long long nq, const KeyMap& km,
                        unsigned int* keys, unsigned long long* vals, hipStream_t s);
int launch_segments(const unsigned int* keys, long long n, long long* seg_begin,
                    long long* seg_end, hipStream_t s);
int launch_plan_positions(const unsigned int* keys, const unsigned long long* vals, long long n,
                          const long long* seg_begin, const unsigned long long* key_base,
                          const int* key_ne, const ClassTab* tab, hipStream_t s);
int launch_far_codes(const ClassParams& cp, long long nq, const int* n64, unsigned char* code,
                     hipStream_t s);
int launch_ket_transpose(const PairListDev& pl, int kmax, const PrimPair* prim, const double* primq,
                         const double* prims, const double* xyz, double* kt, double* kc,
                         hipStream_t s);
int launch_reduce(int n1, int n2, const DestList& dl, const double* tiles, const ShellMap& map,
                  double* X, double* Y, hipStream_t s);
int launch_scale_density(const double* P, const double* fnorm, const MolLayout& ml, double* Ps,
                         hipStream_t s);
int launch_finalize(const double* X, const double* Y, const double* fnorm, const MolLayout& ml,
                    double* J, double* K, hipStream_t s);
int launch_fill(double* p, long long n, double v, hipStream_t s);
int launch_one_electron(int pc, const int* sa, const int* sb, const int* poff, const int* np, int n,
                        const PrimPair* prim, const double* xyz, const int* foff,
                        const double* fnorm, const ShellMap& map, const int* sh_mol,
                        const int* mol_nuc, const double* Z, const double* Cxyz, double* S,
                        double* T, double* V, hipStream_t s);
int launch_eri_components(int nf, const int* npgf, const int* poff, const double* coefs,
                          const double* expns, const double* norms, const int* shell,
                          const double* center, long long nq, const int* quartets, double* out,
                          hipStream_t s);

}  // namespace sf
