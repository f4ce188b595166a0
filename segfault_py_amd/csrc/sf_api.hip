// sf_api.hip -- the C-ABI of libsfeng.so: basis handles (one molecule or a
// batch of independent molecules), Schwarz screening, quartet lists, the
// deterministic digestion plan, the J/K driver with autotuned launch shapes
// and graph replay, one-electron integrals and the nuclear gradient.
//
// The reference's call sites replaced here: rys.eri_rys via f2py
// (src/hf/hf.py:26-30, :100-104), build_eri_parallel (hf.py:33-61) and the
// J/K einsums (hf.py:300).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sfeng.h"
#include "sf_internal.hpp"

namespace sf {

static thread_local char g_err[2048] = "";

void set_error(int code, const char* fmt, ...)
{
    (void)code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t n)
    {
        if (n <= bytes && p) return kOk;
        release();
        if (n == 0) n = 8;
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            set_error(kENoMem, "hipMalloc(%zu) failed", n);
            p = nullptr;
            return kENoMem;
        }
        bytes = n;
        return kOk;
    }
    template <typename T>
    int upload(const std::vector<T>& v)
    {
        int rc = ensure(v.size() * sizeof(T));
        if (rc) return rc;
        if (!v.empty()) SF_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return kOk;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

#define SF_TRY(x)                \
    do {                         \
        int rc_ = (x);           \
        if (rc_ != kOk) return rc_; \
    } while (0)

// Opt-in wall-clock report of setup stages: with SF_SETUP_TIMING set in
// the environment, one line per stage on stderr.
struct StageTimer {
    bool on;
    const char* what;
    std::chrono::steady_clock::time_point t;
    explicit StageTimer(const char* w)
        : on(std::getenv("SF_SETUP_TIMING") != nullptr), what(w), t(std::chrono::steady_clock::now())
    {
    }
    void mark(const char* stage)
    {
        if (!on) return;
        const auto n = std::chrono::steady_clock::now();
        fprintf(stderr, "[sf %s] %s %.3f ms\n", what, stage,
                std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};

// Runs f(lo, hi) over [0, n) split into contiguous ranges on up to 16 host
// threads (setup loops over independent molecules or shell pairs).  Small n
// runs inline.
template <class F>
static void parallel_ranges(long long n, long long min_per_thread, F&& f)
{
    const long long hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long long nt = std::max(1LL, std::min(hw, n / std::max(1LL, min_per_thread)));
    if (nt <= 1) {
        if (n > 0) f(0LL, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (long long t = 1; t < nt; ++t)
        th.emplace_back([&f, t, n, nt] { f(n * t / nt, n * (t + 1) / nt); });
    f(0LL, n / nt);
    for (auto& x : th) x.join();
}

static double dfact2(int n)
{
    double r = 1.0;
    while (n > 1) {
        r *= n;
        n -= 2;
    }
    return r;
}

}  // namespace sf

using namespace sf;

struct PairClassData {
    // every pair of the class, canonical orientation (L(sa) >= L(sb)), both
    // shells in one molecule, molecule by molecule: molecule m's pairs are
    // [mbeg[m], mbeg[m+1])
    std::vector<int> sa, sb, poff, np, mbeg;
    std::vector<double> Q;
    DevBuf d_sa, d_sb, d_poff, d_np, d_Q;
    // significant pairs, molecule by molecule, sorted by Q descending inside
    // each molecule; molecule m's are [s_mbeg[m], s_mbeg[m+1])
    std::vector<int> s_sa, s_sb, s_poff, s_np, s_mol, s_mbeg;
    std::vector<double> s_Q;
    DevBuf ds_sa, ds_sb, ds_poff, ds_np, ds_Q;
    // the significant pairs' primitive pairs transposed (KetView layout,
    // kmax primitives per pair, zero-padded) and their two centres at
    // kc[k * n + j] (Cx Cy Cz Dx Dy Dz), for the kernels' per-lane ket reads
    int kmax = 0;
    DevBuf d_kt, d_kc;
    PairListDev sorted_dev() const
    {
        PairListDev p;
        p.sa = ds_sa.as<int>();
        p.sb = ds_sb.as<int>();
        p.poff = ds_poff.as<int>();
        p.np = ds_np.as<int>();
        p.n = (int)s_sa.size();
        return p;
    }
};

struct QuartetClassData {
    int pcb = 0, pck = 0, same = 0, nbra = 0;
    long long nq_all = 0, nq64 = 0, nq32 = 0;
    // per bra: surviving kets, FP64 kets, first ket and ket count of its molecule
    std::vector<int> n_all, n64, ket0, nket;
    std::vector<long long> base_all, base64, base32;
    DevBuf d_base_all, d_base64, d_base32, d_k032, d_ket0, d_nket, d_pos, d_raw;
    int ntile = 0;
    // one-node classes: class-wide quartet indices of the far quartets
    // (k_far_codes), then of the near ones, of each index space: [0] the
    // whole class, [1] the FP64 segment, [2] the FP32 segment of mixed
    // builds; nlist[sp][0] near and [1] far quartets
    bool lists = false;
    DevBuf d_list[3];
    long long nlist[3][2] = {};
};

// A captured J/K build (sf_jk_graph) and the arguments it was captured for.
struct JKGraph {
    hipGraphExec_t exec = nullptr;
    const double* P = nullptr;
    double* J = nullptr;
    double* K = nullptr;
    int precision = -1;
};

struct sf_basis {
    int nshell = 0, nbf = 0, natom = 0, nmol = 1;
    bool have_shell_atom = false;          // shell_atom given at creation (sf_grad needs it)
    std::vector<int> L, nprim, pstart, foff, atom;
    std::vector<int> sh_mol, sh_loc;       // molecule of each shell, index inside it
    std::vector<int> mol_shell, mol_atom;  // nmol+1 ranges of shells and of nuclei
    std::vector<int> mol_nbf, mol_fbase;   // functions per molecule, first global function
    std::vector<long long> mol_off;        // layout: first element of each molecule's block
    std::vector<int> mol_ld;               //         and its leading dimension
    long long extent = 0;                  // elements spanned by the layout
    std::vector<double> xyz, expn, coef, fnorm;
    std::vector<PrimPair> prim;
    std::vector<double> prim_q;            // Schwarz factor of each primitive pair
    std::vector<double> prim_s;            // pi^{3/2} K p^{-3/2} of each primitive pair
    DevBuf d_prim, d_xyz, d_foff, d_fnorm, d_zero, d_prim_q, d_prim_s;
    DevBuf d_row, d_col, d_ld, d_sh_mol, d_sh_loc, d_mol_ns, d_mol_kbase;
    DevBuf d_mol_off, d_mol_ld, d_mol_n, d_mol_fbase, d_mol_nuc;
    // nuclei (point charges) for the one-electron potential
    int nnuc = 0;
    DevBuf d_Z, d_nuc, d_atom;
    // nuclear gradient scratch (built on first use after each sf_screen)
    bool grad_ready = false;
    // how a class's FP32 derivative quartets run in mixed gradients: -1 =
    // not yet measured, 0 = their own launch in FP32 arithmetic, 1 = one
    // launch over the class in FP64 arithmetic (chosen by timing both on
    // the first mixed sf_grad after each sf_screen)
    int grad_arith[kNumQuartetClass];
    std::vector<long long> gpart_bra, gpart_ket;  // per class offsets into d_gpart
    DevBuf d_gpart, d_gout, d_gcsr_off, d_gcsr_idx, d_grows, d_Ws, d_g2, d_gtot;
    DevBuf d_gonee;  // one-electron gradient: per (pair, nucleus) d/dA, d/dB partials
    DevBuf d_gkc;    // per quartet of one class: ket, then bra derivatives summed over slices
    long long gkc_q = 1;  // quartets d_gkc holds per half
    PairClassData pc[kNumPairClass];
    std::vector<int> pair_slot;  // one molecule: pair (a>=b) -> (pc << 24) | position
    bool have_schwarz = false, screened = false;
    double tau = 0.0, tau_mixed = 0.0;
    // primitive-quartet thresholds.  FP32 quartets: tau_prim_set < 0 means
    // 1e-8 * tau_mixed (about FP32 resolution of the largest FP32 quartet),
    // else the value set by sf_set_prim_tau.  FP64 quartets: tau_prim64_set
    // < 0 means 1e-3 * tau (the dropped primitive terms of a quartet stay
    // below the shell-level Schwarz threshold), else sf_set_prim_tau_fp64's.
    double tau_prim_set = -1.0, tau_prim = 0.0;
    double tau_prim64_set = -1.0, tau_prim64 = 0.0;
    QuartetClassData qc[kNumQuartetClass];
    long long funcq = 0, primq = 0;
    // digestion plan: destination-sorted tile slots
    DevBuf d_segb, d_sege, d_key_base, d_tiles;
    // reduction lists per tile shape g = 3 L1 + L2 (see DestList)
    DevBuf d_ds1[kNumTileGroup], d_ds2[kNumTileGroup], d_dxy[kNumTileGroup],
        d_ch_src[kNumTileGroup], d_ch_n[kNumTileGroup], d_ch_dest[kNumTileGroup],
        d_ch_part[kNumTileGroup];
    DevBuf d_m_dest[kNumTileGroup], d_m_pbase[kNumTileGroup], d_m_nch[kNumTileGroup], d_part;
    int ndest[kNumTileGroup] = {0}, nchunk[kNumTileGroup] = {0}, nmulti[kNumTileGroup] = {0};
    long long part_off[kNumTileGroup] = {0};
    unsigned long long ntile_total = 0;
    DevBuf d_Ps, d_X, d_Y;
    // launch shapes per quartet class and precision segment (sf_autotune)
    // launch shapes per class and tuning slot (tune_slot: the whole class in
    // FP64 builds; the FP64 and the FP32 segment of mixed builds)
    Tune tune[kNumQuartetClass][3];
    // integral engine per quartet class: 0 = Rys quadrature, 1 = McMurchie-
    // Davidson with MFMA transforms (FP64 in both precision segments)
    int engine[kNumQuartetClass] = {0};
    JKGraph graph;
    hipStream_t cap_stream = nullptr;
    // concurrent builds: the class launches and the per-shape reductions of
    // sf_jk / sf_jk_graph are spread over the caller's stream and `nside`
    // (eager builds) / `nside_graph` (captured builds) side streams (0 = all
    // on the caller's stream), forked from and joined back into the
    // caller's stream with events.  Measured on 20 waters (FP64 / mixed ms
    // per build, eager | graph): one side stream 6.27 / 6.16 | 6.25 / 6.13,
    // two 6.04 / 5.92 | 6.41 / 6.39, three 6.26 / 6.03 | 6.11 / 5.92 (no
    // side stream: 6.74 FP64 eager).
    static constexpr int kMaxSide = 3;
    int nside = 2;
    int nside_graph = 3;
    // per class and precision, its launches' time in one build (ms; 0 =
    // unknown), from sf_autotune's estimates or a sf_jk_profile build: the
    // cost the launches are balanced over the streams by
    double class_ms[kNumQuartetClass][2] = {};
    hipStream_t side[kMaxSide] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxSide] = {};

    ~sf_basis()
    {
        drop_graph();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (int k = 0; k < kMaxSide; ++k) {
            if (side[k]) (void)hipStreamDestroy(side[k]);
            if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
    }
    int ensure_side()
    {
        // every side stream either build path may use (nside, nside_graph)
        for (int k = 0; k < std::max(nside, nside_graph); ++k) {
            if (!side[k]) SF_HIP(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking));
            if (!ev_join[k]) SF_HIP(hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming));
        }
        if (!ev_fork) SF_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        return kOk;
    }
    void drop_graph()
    {
        if (graph.exec) (void)hipGraphExecDestroy(graph.exec);
        graph = JKGraph();
    }
    ShellMap map() const
    {
        ShellMap m;
        m.row = d_row.as<long long>();
        m.col = d_col.as<int>();
        m.ld = d_ld.as<int>();
        return m;
    }
    MolLayout layout() const
    {
        MolLayout ml;
        ml.off = d_mol_off.as<long long>();
        ml.ld = d_mol_ld.as<int>();
        ml.n = d_mol_n.as<int>();
        ml.fbase = d_mol_fbase.as<int>();
        ml.nmol = nmol;
        ml.nmax = *std::max_element(mol_nbf.begin(), mol_nbf.end());
        return ml;
    }
};

extern "C" const char* sf_last_error(void) { return g_err; }
extern "C" int sf_version(void) { return 200; }

// ------------------------------------------------------------ batches ----
extern "C" int sf_boys(long long n, int mmax, const double* T, double* F, void* stream)
{
    if (n < 0 || (n > 0 && (!T || !F))) {
        set_error(kEInval, "sf_boys: bad buffers");
        return kEInval;
    }
    return launch_boys(n, mmax, T, F, (hipStream_t)stream);
}

extern "C" int sf_rys_roots(long long n, int nroots, const double* X, double* u, double* w,
                            void* stream)
{
    if (n < 0 || (n > 0 && (!X || !u || !w))) {
        set_error(kEInval, "sf_rys_roots: bad buffers");
        return kEInval;
    }
    return launch_rys(n, nroots, X, u, w, (hipStream_t)stream);
}

// Argument checks shared by sf_eri_components and its allocation-free
// variant; *nprim_out = primitives the function table spans.
static int check_components_args(int nf, const int* npgf, const int* poff, const double* coefs,
                                 const double* expns, const double* norms, const int* shell,
                                 const double* center, long long nq, const int* quartets,
                                 double* out, long long* nprim_out)
{
    if (nf <= 0 || nq < 0 || !npgf || !poff || !coefs || !expns || !norms || !shell || !center ||
        (nq > 0 && (!quartets || !out))) {
        set_error(kEInval, "sf_eri_components: bad arguments");
        return kEInval;
    }
    long long nprim = 0;
    for (int f = 0; f < nf; ++f) {
        if (npgf[f] <= 0 || poff[f] < 0) {
            set_error(kEInval, "function %d: npgf %d poff %d", f, npgf[f], poff[f]);
            return kEInval;
        }
        nprim = std::max<long long>(nprim, (long long)poff[f] + npgf[f]);
        for (int e = 0; e < 3; ++e)
            if (shell[3 * f + e] < 0) {
                set_error(kEInval, "function %d: negative angular exponent", f);
                return kEInval;
            }
    }
    for (long long q = 0; q < nq; ++q) {
        int L = 0;
        for (int k = 0; k < 4; ++k) {
            const int f = quartets[4 * q + k];
            if (f < 0 || f >= nf) {
                set_error(kEInval, "quartet %lld: function index %d out of range", q, f);
                return kEInval;
            }
            L += shell[3 * f] + shell[3 * f + 1] + shell[3 * f + 2];
        }
        if (L > 13) {
            set_error(kEUnsupported, "quartet %lld: total angular momentum %d > 13", q, L);
            return kEUnsupported;
        }
    }
    *nprim_out = nprim;
    return kOk;
}

extern "C" int sf_eri_components(int nf, const int* npgf, const int* poff, const double* coefs,
                                 const double* expns, const double* norms, const int* shell,
                                 const double* center, long long nq, const int* quartets,
                                 double* out)
{
    long long nprim = 0;
    SF_TRY(check_components_args(nf, npgf, poff, coefs, expns, norms, shell, center, nq, quartets,
                                 out, &nprim));
    if (nq == 0) return kOk;
    DevBuf a, b, c, d, e, f, g, h, qb, ob;
    SF_TRY(a.ensure(nf * sizeof(int)));
    SF_TRY(b.ensure(nf * sizeof(int)));
    SF_TRY(c.ensure(nprim * sizeof(double)));
    SF_TRY(d.ensure(nprim * sizeof(double)));
    SF_TRY(e.ensure(nprim * sizeof(double)));
    SF_TRY(g.ensure(3 * nf * sizeof(int)));
    SF_TRY(h.ensure(3 * nf * sizeof(double)));
    SF_TRY(qb.ensure(4 * nq * sizeof(int)));
    SF_TRY(ob.ensure(nq * sizeof(double)));
    SF_HIP(hipMemcpy(a.p, npgf, nf * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(b.p, poff, nf * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(c.p, coefs, nprim * sizeof(double), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(d.p, expns, nprim * sizeof(double), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(e.p, norms, nprim * sizeof(double), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(g.p, shell, 3 * nf * sizeof(int), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(h.p, center, 3 * nf * sizeof(double), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(qb.p, quartets, 4 * nq * sizeof(int), hipMemcpyHostToDevice));
    SF_TRY(launch_eri_components(nf, a.as<int>(), b.as<int>(), c.as<double>(), d.as<double>(),
                                 e.as<double>(), g.as<int>(), h.as<double>(), nq, qb.as<int>(),
                                 ob.as<double>(), 0));
    SF_HIP(hipMemcpy(out, ob.p, nq * sizeof(double), hipMemcpyDeviceToHost));
    return kOk;
}

// ---------------------------------------------- allocation-free staging ----
// sf_eri_components_noalloc stages the function table, the quartet list
// and the results in this array instead of allocating device memory per
// call.  In host code `scratch` names a host-side stand-in, not the device
// address: the address is fetched once per device with hipGetSymbolAddress
// (scratch_ptr) and only that pointer is handed to copies and kernels.
// Every device holds its own copy until the process ends; calls on one
// device share it, so calls that overlap on one device race.
static constexpr size_t kScratchBytes = 32u << 20;  // 32 MiB per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxScratchDevices = 64;
static void* g_scratch_dev[kMaxScratchDevices] = {nullptr};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up on the
// first call made with that device current).
static int scratch_ptr(unsigned char** out)
{
    int dev = 0;
    SF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_error(kEUnsupported, "device %d: beyond the %d devices with staging scratch", dev,
                  kMaxScratchDevices);
        return kEUnsupported;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        SF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error(kEHip, "hipGetSymbolAddress(scratch) returned a null address");
            return kEHip;
        }
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_dev[dev]);
    return kOk;
}

static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

extern "C" int sf_eri_components_noalloc(int nf, const int* npgf, const int* poff,
                                         const double* coefs, const double* expns,
                                         const double* norms, const int* shell,
                                         const double* center, long long nq, const int* quartets,
                                         double* out)
{
    long long nprim = 0;
    SF_TRY(check_components_args(nf, npgf, poff, coefs, expns, norms, shell, center, nq, quartets,
                                 out, &nprim));
    if (nq == 0) return kOk;
    // sections of the staging array, each 256-byte aligned
    const size_t sz[9] = {nf * sizeof(int),          nf * sizeof(int),
                          nprim * sizeof(double),    nprim * sizeof(double),
                          nprim * sizeof(double),    3 * (size_t)nf * sizeof(int),
                          3 * (size_t)nf * sizeof(double), 4 * (size_t)nq * sizeof(int),
                          (size_t)nq * sizeof(double)};
    size_t off[9], total = 0;
    for (int k = 0; k < 9; ++k) {
        off[k] = total;
        total += align256(sz[k]);
    }
    if (total > kScratchBytes) {
        set_error(kEUnsupported,
                  "sf_eri_components_noalloc: %zu bytes of tables and results exceed the %zu-byte "
                  "staging array (use sf_eri_components, or fewer quartets per call)",
                  total, kScratchBytes);
        return kEUnsupported;
    }
    unsigned char* base = nullptr;
    SF_TRY(scratch_ptr(&base));
    const void* src[8] = {npgf, poff, coefs, expns, norms, shell, center, quartets};
    for (int k = 0; k < 8; ++k)
        SF_HIP(hipMemcpy(base + off[k], src[k], sz[k], hipMemcpyHostToDevice));
    SF_TRY(launch_eri_components(
        nf, reinterpret_cast<const int*>(base + off[0]), reinterpret_cast<const int*>(base + off[1]),
        reinterpret_cast<const double*>(base + off[2]), reinterpret_cast<const double*>(base + off[3]),
        reinterpret_cast<const double*>(base + off[4]), reinterpret_cast<const int*>(base + off[5]),
        reinterpret_cast<const double*>(base + off[6]), nq,
        reinterpret_cast<const int*>(base + off[7]), reinterpret_cast<double*>(base + off[8]), 0));
    SF_HIP(hipMemcpy(out, base + off[8], sz[8], hipMemcpyDeviceToHost));
    return kOk;
}

extern "C" int sf_eri_rys(const int npgf[4], const double* coefs, const double* expns,
                          const double* norms, const int shells[12], const double centers[12],
                          double* out)
{
    if (!npgf || !coefs || !expns || !norms || !shells || !centers || !out) {
        set_error(kEInval, "sf_eri_rys: null argument");
        return kEInval;
    }
    int poff[4] = {0, npgf[0], npgf[0] + npgf[1], npgf[0] + npgf[1] + npgf[2]};
    const int q[4] = {0, 1, 2, 3};
    return sf_eri_components(4, npgf, poff, coefs, expns, norms, shells, centers, 1, q, out);
}

// -------------------------------------------------------------- basis ----
// Per-shell matrix positions and per-molecule layout arrays on the device,
// internal buffers sized to the layout.  Called after creation and after
// every sf_basis_set_layout.
static int apply_layout(sf_basis* b)
{
    std::vector<long long> row(b->nshell);
    std::vector<int> col(b->nshell), ld(b->nshell);
    long long extent = 0;
    for (int m = 0; m < b->nmol; ++m)
        extent = std::max(extent, b->mol_off[m] + (long long)b->mol_nbf[m] * b->mol_ld[m]);
    for (int s = 0; s < b->nshell; ++s) {
        const int m = b->sh_mol[s];
        const int lf = b->foff[s] - b->mol_fbase[m];
        row[s] = b->mol_off[m] + (long long)lf * b->mol_ld[m];
        col[s] = lf;
        ld[s] = b->mol_ld[m];
    }
    b->extent = extent;
    SF_TRY(b->d_row.upload(row));
    SF_TRY(b->d_col.upload(col));
    SF_TRY(b->d_ld.upload(ld));
    SF_TRY(b->d_mol_off.upload(b->mol_off));
    SF_TRY(b->d_mol_ld.upload(b->mol_ld));
    SF_TRY(b->d_Ps.ensure((size_t)std::max(1LL, extent) * sizeof(double)));
    SF_TRY(b->d_X.ensure((size_t)std::max(1LL, extent) * sizeof(double)));
    SF_TRY(b->d_Y.ensure((size_t)std::max(1LL, extent) * sizeof(double)));
    SF_TRY(b->d_Ws.ensure((size_t)std::max(1LL, extent) * sizeof(double)));
    b->drop_graph();
    return kOk;
}

// Shared by sf_basis_create (one molecule) and sf_basis_create_batch.
static int create_basis(int nmol, const int* mol_shell, int nshell, const int* L, const int* nprim,
                        const double* center, const double* expns, const double* coefs,
                        const int* mol_atom, int natom, const int* shell_atom, sf_basis** out)
{
    if (!out || nshell <= 0 || nmol <= 0 || !mol_shell || !L || !nprim || !center || !expns ||
        !coefs || natom < 0) {
        set_error(kEInval, "sf_basis_create: bad arguments");
        return kEInval;
    }
    *out = nullptr;
    if (mol_shell[0] != 0 || mol_shell[nmol] != nshell) {
        set_error(kEInval, "molecule shell ranges must start at 0 and end at nshell");
        return kEInval;
    }
    for (int m = 0; m < nmol; ++m)
        if (mol_shell[m + 1] <= mol_shell[m]) {
            set_error(kEInval, "molecule %d has no shells", m);
            return kEInval;
        }
    if (mol_atom) {
        if (mol_atom[0] < 0 || mol_atom[nmol] > natom) {
            set_error(kEInval, "molecule atom ranges outside [0, %d]", natom);
            return kEInval;
        }
        for (int m = 0; m < nmol; ++m)
            if (mol_atom[m + 1] < mol_atom[m]) {
                set_error(kEInval, "molecule %d: decreasing atom range", m);
                return kEInval;
            }
    }
    StageTimer tm("create");
    sf_basis* b = new (std::nothrow) sf_basis();
    if (!b) {
        set_error(kENoMem, "sf_basis_create: host allocation");
        return kENoMem;
    }
    b->nshell = nshell;
    b->natom = natom;
    b->nmol = nmol;
    b->have_shell_atom = shell_atom != nullptr;
    b->mol_shell.assign(mol_shell, mol_shell + nmol + 1);
    if (mol_atom) b->mol_atom.assign(mol_atom, mol_atom + nmol + 1);
    else b->mol_atom = {0, natom};
    int off = 0, pst = 0;
    for (int m = 0; m < nmol; ++m) {
        b->mol_fbase.push_back(off);
        for (int s = mol_shell[m]; s < mol_shell[m + 1]; ++s) {
            if (L[s] < 0 || L[s] > kMaxL) {
                set_error(kEUnsupported, "shell %d: L = %d (s, p, d, f supported)", s, L[s]);
                delete b;
                return kEUnsupported;
            }
            if (nprim[s] <= 0 || nprim[s] > 64) {
                set_error(kEInval, "shell %d: nprim = %d", s, nprim[s]);
                delete b;
                return kEInval;
            }
            // the atom of a shell must be one of its own molecule's nuclei:
            // the gradient sums per atom through it (host lists and device
            // rows), so an index outside the range would address past them
            if (shell_atom && (shell_atom[s] < b->mol_atom[m] || shell_atom[s] >= b->mol_atom[m + 1])) {
                set_error(kEInval, "shell %d: atom %d outside molecule %d's atoms [%d, %d)", s,
                          shell_atom[s], m, b->mol_atom[m], b->mol_atom[m + 1]);
                delete b;
                return kEInval;
            }
            b->L.push_back(L[s]);
            b->nprim.push_back(nprim[s]);
            b->pstart.push_back(pst);
            b->foff.push_back(off);
            b->atom.push_back(shell_atom ? shell_atom[s] : 0);
            b->sh_mol.push_back(m);
            b->sh_loc.push_back(s - mol_shell[m]);
            for (int k = 0; k < 3; ++k) {
                if (!std::isfinite(center[3 * s + k])) {
                    set_error(kEInval, "shell %d: centre is not finite", s);
                    delete b;
                    return kEInval;
                }
                b->xyz.push_back(center[3 * s + k]);
            }
            for (int i = 0; i < nprim[s]; ++i) {
                if (!(expns[pst + i] > 0.0) || !std::isfinite(expns[pst + i]) ||
                    !std::isfinite(coefs[pst + i])) {
                    set_error(kEInval, "shell %d: primitive %d has exponent %g, coefficient %g", s,
                              i, expns[pst + i], coefs[pst + i]);
                    delete b;
                    return kEInval;
                }
                b->expn.push_back(expns[pst + i]);
                b->coef.push_back(coefs[pst + i]);
            }
            // per-component angular factors, reference component order
            for (int lx = L[s]; lx >= 0; --lx)
                for (int ly = L[s] - lx; ly >= 0; --ly) {
                    const int lz = L[s] - lx - ly;
                    b->fnorm.push_back(1.0 / std::sqrt(dfact2(2 * lx - 1) * dfact2(2 * ly - 1) *
                                                       dfact2(2 * lz - 1)));
                }
            off += (L[s] + 1) * (L[s] + 2) / 2;
            pst += nprim[s];
        }
        b->mol_nbf.push_back(off - b->mol_fbase[m]);
    }
    b->nbf = off;
    tm.mark("shells");
    // plan keys: 2 ns^2 per molecule (X and Y destinations)
    std::vector<unsigned> kbase(nmol);
    std::vector<int> mol_ns(nmol);
    unsigned long long kacc = 0;
    for (int m = 0; m < nmol; ++m) {
        mol_ns[m] = mol_shell[m + 1] - mol_shell[m];
        kbase[m] = (unsigned)kacc;
        kacc += 2ULL * mol_ns[m] * mol_ns[m];
        if (kacc >= (1ULL << 31)) {
            set_error(kEUnsupported, "too many shell pairs for the digestion plan");
            delete b;
            return kEUnsupported;
        }
    }
    // Shell pairs of each molecule, canonical orientation, primitive pair
    // data.  Primitive-pair screening: with Kp = K/p, a primitive quartet is
    // bounded by 2 pi^{5/2} |Kp_ab| |Kp_cd| / sqrt(p+q) times its polynomial
    // factors; pairs with 2 pi^{5/2} |Kp_ab| M / sqrt(p_ab) < 1e-20 (M =
    // largest |Kp| of the molecule) are dropped, which changes no integral by
    // more than ~1e-15.
    const double kPrimTol = 1e-20;
    // molecules are independent: each fills its own lists on a host thread,
    // then the lists are concatenated in molecule order
    struct MolPairs {
        std::vector<int> sa[kNumPairClass], sb[kNumPairClass], poff[kNumPairClass], np[kNumPairClass];
        std::vector<PrimPair> prim;
    };
    std::vector<MolPairs> mp(nmol);
    if (nmol == 1) b->pair_slot.resize((size_t)nshell * (nshell + 1) / 2);
    const sf_basis* cb = b;
    parallel_ranges(nmol, 4, [&](long long m0, long long m1) {
        for (long long m = m0; m < m1; ++m) {
            MolPairs& R = mp[m];
            const int s0 = mol_shell[m], s1 = mol_shell[m + 1];
            double kp_max = 0.0;
            for (int a = s0; a < s1; ++a)
                for (int c = s0; c <= a; ++c) {
                    const double* A = &cb->xyz[3 * a];
                    const double* B = &cb->xyz[3 * c];
                    const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                                       (A[2] - B[2]) * (A[2] - B[2]);
                    for (int i = 0; i < nprim[a]; ++i)
                        for (int j = 0; j < nprim[c]; ++j) {
                            const double ea = cb->expn[cb->pstart[a] + i], eb = cb->expn[cb->pstart[c] + j];
                            const double p = ea + eb;
                            const double K = cb->coef[cb->pstart[a] + i] * cb->coef[cb->pstart[c] + j] *
                                             std::exp(-ea * eb / p * AB2);
                            kp_max = std::max(kp_max, std::fabs(K / p));
                        }
                }
            for (int a = s0; a < s1; ++a)
                for (int c = s0; c <= a; ++c) {
                    int sa = a, sb = c;
                    if (L[sb] > L[sa]) std::swap(sa, sb);
                    const int pcl = pair_class(L[sa], L[sb]);
                    R.sa[pcl].push_back(sa);
                    R.sb[pcl].push_back(sb);
                    R.poff[pcl].push_back((int)R.prim.size());
                    if (nmol == 1)
                        b->pair_slot[(size_t)a * (a + 1) / 2 + c] = (pcl << 24) | (int)(R.sa[pcl].size() - 1);
                    const double* A = &cb->xyz[3 * sa];
                    const double* B = &cb->xyz[3 * sb];
                    const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                                       (A[2] - B[2]) * (A[2] - B[2]);
                    int kept = 0;
                    for (int i = 0; i < nprim[sa]; ++i)
                        for (int j = 0; j < nprim[sb]; ++j) {
                            const double ea = cb->expn[cb->pstart[sa] + i], eb = cb->expn[cb->pstart[sb] + j];
                            const double p = ea + eb;
                            PrimPair pp;
                            pp.p = p;
                            pp.Px = (ea * A[0] + eb * B[0]) / p;
                            pp.Py = (ea * A[1] + eb * B[1]) / p;
                            pp.Pz = (ea * A[2] + eb * B[2]) / p;
                            pp.K = cb->coef[cb->pstart[sa] + i] * cb->coef[cb->pstart[sb] + j] *
                                   std::exp(-ea * eb / p * AB2);
                            pp.ip = 1.0 / p;
                            pp.Kp = pp.K / p;
                            pp.eb = eb;
                            if (34.986836655249725 * std::fabs(pp.Kp) * kp_max / std::sqrt(p) < kPrimTol)
                                continue;
                            R.prim.push_back(pp);
                            ++kept;
                        }
                    R.np[pcl].push_back(kept);
                }
        }
    });
    {
        size_t nprim_all = 0;
        for (const auto& R : mp) nprim_all += R.prim.size();
        if (nprim_all >= (1u << 31)) {
            set_error(kEUnsupported, "too many primitive pairs");
            delete b;
            return kEUnsupported;
        }
        b->prim.reserve(nprim_all);
    }
    for (auto& P : b->pc) P.mbeg.push_back(0);
    for (int m = 0; m < nmol; ++m) {
        MolPairs& R = mp[m];
        const int base = (int)b->prim.size();
        b->prim.insert(b->prim.end(), R.prim.begin(), R.prim.end());
        for (int p = 0; p < kNumPairClass; ++p) {
            PairClassData& P = b->pc[p];
            P.sa.insert(P.sa.end(), R.sa[p].begin(), R.sa[p].end());
            P.sb.insert(P.sb.end(), R.sb[p].begin(), R.sb[p].end());
            P.np.insert(P.np.end(), R.np[p].begin(), R.np[p].end());
            for (int o : R.poff[p]) P.poff.push_back(base + o);
            P.mbeg.push_back((int)P.sa.size());
        }
        std::vector<PrimPair>().swap(R.prim);
    }
    tm.mark("primitive pairs");
    // default layout: molecule blocks packed back to back, ld = n
    long long lo = 0;
    for (int m = 0; m < nmol; ++m) {
        b->mol_off.push_back(lo);
        b->mol_ld.push_back(b->mol_nbf[m]);
        lo += (long long)b->mol_nbf[m] * b->mol_nbf[m];
    }
    int rc = b->d_prim.upload(b->prim);
    if (!rc) rc = b->d_atom.upload(b->atom);
    if (!rc) rc = b->d_xyz.upload(b->xyz);
    if (!rc) rc = b->d_foff.upload(b->foff);
    if (!rc) rc = b->d_fnorm.upload(b->fnorm);
    if (!rc) rc = b->d_sh_mol.upload(b->sh_mol);
    if (!rc) rc = b->d_sh_loc.upload(b->sh_loc);
    if (!rc) rc = b->d_mol_ns.upload(mol_ns);
    if (!rc) rc = b->d_mol_kbase.upload(kbase);
    if (!rc) rc = b->d_mol_n.upload(b->mol_nbf);
    if (!rc) rc = b->d_mol_fbase.upload(b->mol_fbase);
    if (!rc) rc = b->d_mol_nuc.upload(b->mol_atom);
    for (int p = 0; p < kNumPairClass && !rc; ++p) {
        PairClassData& P = b->pc[p];
        rc = P.d_sa.upload(P.sa);
        if (!rc) rc = P.d_sb.upload(P.sb);
        if (!rc) rc = P.d_poff.upload(P.poff);
        if (!rc) rc = P.d_np.upload(P.np);
        if (!rc) rc = P.d_Q.ensure(std::max<size_t>(1, P.sa.size()) * sizeof(double));
    }
    if (!rc) rc = apply_layout(b);
    if (rc) {
        delete b;
        return rc;
    }
    tm.mark("uploads");
    *out = b;
    return kOk;
}

extern "C" int sf_basis_create(int nshell, const int* L, const int* nprim, const double* center,
                               const double* expns, const double* coefs, int natom,
                               const int* shell_atom, sf_basis** out)
{
    const int ms[2] = {0, nshell};
    const int ma[2] = {0, natom < 0 ? 0 : natom};
    return create_basis(1, ms, nshell, L, nprim, center, expns, coefs, ma, natom, shell_atom, out);
}

extern "C" int sf_basis_create_batch(int nmol, const int* mol_shell_begin, int nshell, const int* L,
                                     const int* nprim, const double* center, const double* expns,
                                     const double* coefs, int natom, const int* mol_atom_begin,
                                     const int* shell_atom, sf_basis** out)
{
    if (!mol_atom_begin) {
        set_error(kEInval, "sf_basis_create_batch: mol_atom_begin is required");
        return kEInval;
    }
    return create_basis(nmol, mol_shell_begin, nshell, L, nprim, center, expns, coefs,
                        mol_atom_begin, natom, shell_atom, out);
}

extern "C" int sf_basis_set_layout(sf_basis* b, const long long* block_off, const int* block_ld)
{
    if (!b || !block_off || !block_ld) {
        set_error(kEInval, "sf_basis_set_layout: null argument");
        return kEInval;
    }
    std::vector<std::pair<long long, long long>> span;
    for (int m = 0; m < b->nmol; ++m) {
        if (block_off[m] < 0 || block_ld[m] < b->mol_nbf[m]) {
            set_error(kEInval, "molecule %d: offset %lld, ld %d (needs ld >= %d)", m, block_off[m],
                      block_ld[m], b->mol_nbf[m]);
            return kEInval;
        }
        span.emplace_back(block_off[m], block_off[m] + (long long)b->mol_nbf[m] * block_ld[m]);
    }
    std::sort(span.begin(), span.end());
    for (size_t k = 1; k < span.size(); ++k)
        if (span[k].first < span[k - 1].second) {
            set_error(kEInval, "molecule blocks overlap in the layout");
            return kEInval;
        }
    b->mol_off.assign(block_off, block_off + b->nmol);
    b->mol_ld.assign(block_ld, block_ld + b->nmol);
    return apply_layout(b);
}

extern "C" int sf_basis_layout(const sf_basis* b, int* nmol, int* mol_nbf, long long* block_off,
                               int* block_ld, long long* extent)
{
    if (!b || !nmol) {
        set_error(kEInval, "sf_basis_layout: null argument");
        return kEInval;
    }
    *nmol = b->nmol;
    for (int m = 0; m < b->nmol; ++m) {
        if (mol_nbf) mol_nbf[m] = b->mol_nbf[m];
        if (block_off) block_off[m] = b->mol_off[m];
        if (block_ld) block_ld[m] = b->mol_ld[m];
    }
    if (extent) *extent = b->extent;
    return kOk;
}

extern "C" int sf_basis_destroy(sf_basis* b)
{
    delete b;
    return kOk;
}

extern "C" int sf_basis_nbf(const sf_basis* b, int* nbf)
{
    if (!b || !nbf) {
        set_error(kEInval, "sf_basis_nbf: null argument");
        return kEInval;
    }
    *nbf = b->nbf;
    return kOk;
}

// Schwarz factors of every shell pair and of every primitive pair; then the
// primitive pairs of each shell pair are reordered by their own factor,
// largest first, so that screened primitive loops can stop early.
static int compute_schwarz(sf_basis* b)
{
    if (b->have_schwarz) return kOk;
    SF_TRY(b->d_prim_q.ensure(std::max<size_t>(1, b->prim.size()) * sizeof(double)));
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        const int n = (int)P.sa.size();
        if (!n) continue;
        SF_TRY(launch_schwarz(p, P.d_sa.as<int>(), P.d_sb.as<int>(), P.d_poff.as<int>(),
                              P.d_np.as<int>(), n, b->d_prim.as<PrimPair>(), b->d_xyz.as<double>(),
                              P.d_Q.as<double>(), b->d_prim_q.as<double>(), 0));
    }
    SF_HIP(hipDeviceSynchronize());
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        P.Q.resize(P.sa.size());
        if (!P.Q.empty())
            SF_HIP(hipMemcpy(P.Q.data(), P.d_Q.p, P.Q.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    b->prim_q.resize(b->prim.size());
    if (!b->prim.empty())
        SF_HIP(hipMemcpy(b->prim_q.data(), b->d_prim_q.p, b->prim_q.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
    // shell pairs own disjoint primitive ranges: sorted on several threads
    for (auto& P : b->pc)
        parallel_ranges((long long)P.sa.size(), 256, [&](long long i0, long long i1) {
            std::vector<int> order;
            std::vector<PrimPair> pp;
            std::vector<double> qq;
            for (long long i = i0; i < i1; ++i) {
                const int o = P.poff[i], n = P.np[i];
                order.resize(n);
                std::iota(order.begin(), order.end(), 0);
                std::stable_sort(order.begin(), order.end(),
                                 [&](int x, int y) { return b->prim_q[o + x] > b->prim_q[o + y]; });
                pp.assign(b->prim.begin() + o, b->prim.begin() + o + n);
                qq.assign(b->prim_q.begin() + o, b->prim_q.begin() + o + n);
                for (int k = 0; k < n; ++k) {
                    b->prim[o + k] = pp[order[k]];
                    b->prim_q[o + k] = qq[order[k]];
                }
            }
        });
    SF_TRY(b->d_prim.upload(b->prim));
    SF_TRY(b->d_prim_q.upload(b->prim_q.empty() ? std::vector<double>(1, 0.0) : b->prim_q));
    // charge of each primitive pair for the far one-node quartets
    // (eri_one_node_far): S = pi^{3/2} K p^{-3/2} = pi^{3/2} (K/p) sqrt(1/p)
    b->prim_s.resize(b->prim.size());
    for (size_t k = 0; k < b->prim.size(); ++k)
        b->prim_s[k] = 5.568327996831708 * b->prim[k].Kp * std::sqrt(b->prim[k].ip);
    SF_TRY(b->d_prim_s.upload(b->prim_s.empty() ? std::vector<double>(1, 0.0) : b->prim_s));
    b->have_schwarz = true;
    return kOk;
}

extern "C" int sf_schwarz(sf_basis* b, double* Q)
{
    if (!b || !Q) {
        set_error(kEInval, "sf_schwarz: null argument");
        return kEInval;
    }
    if (b->nmol != 1) {
        set_error(kEUnsupported, "sf_schwarz: one molecule only (use sf_pair_list for a batch)");
        return kEUnsupported;
    }
    SF_TRY(compute_schwarz(b));
    for (size_t k = 0; k < b->pair_slot.size(); ++k) {
        const int s = b->pair_slot[k];
        Q[k] = b->pc[s >> 24].Q[s & 0xffffff];
    }
    return kOk;
}

static int key_bits(unsigned long long nkeys)
{
    int bits = 1;
    while ((1ULL << bits) < nkeys) ++bits;
    return bits;
}

static int build_far_lists(sf_basis* b);

extern "C" int sf_screen(sf_basis* b, double tau, double tau_mixed)
{
    if (!b || !(tau >= 0.0) || !(tau_mixed >= 0.0)) {
        set_error(kEInval, "sf_screen: bad arguments");
        return kEInval;
    }
    StageTimer tm("screen");
    SF_TRY(compute_schwarz(b));
    tm.mark("schwarz");
    b->drop_graph();
    b->screened = false;
    b->grad_ready = false;
    for (int c = 0; c < kNumQuartetClass; ++c) b->grad_arith[c] = -1;
    b->tau = tau;
    b->tau_mixed = tau_mixed;
    b->tau_prim = b->tau_prim_set >= 0.0 ? b->tau_prim_set : 1e-8 * tau_mixed;
    b->tau_prim64 = b->tau_prim64_set >= 0.0 ? b->tau_prim64_set : 1e-3 * tau;
    const int nmol = b->nmol;
    // largest Schwarz factor of each molecule
    std::vector<double> qmax(nmol, 0.0);
    for (auto& P : b->pc)
        for (int m = 0; m < nmol; ++m)
            for (int i = P.mbeg[m]; i < P.mbeg[m + 1]; ++i) qmax[m] = std::max(qmax[m], P.Q[i]);
    // significant pairs of each molecule, sorted by Q descending (ties: original order)
    std::vector<std::vector<int>> sig(nmol);
    for (auto& P : b->pc) {
        P.s_sa.clear(); P.s_sb.clear(); P.s_poff.clear(); P.s_np.clear(); P.s_Q.clear();
        P.s_mol.clear(); P.s_mbeg.assign(1, 0);
        parallel_ranges(nmol, 8, [&](long long m0, long long m1) {
            for (long long m = m0; m < m1; ++m) {
                std::vector<int>& idx = sig[m];
                idx.clear();
                for (int i = P.mbeg[m]; i < P.mbeg[m + 1]; ++i)
                    if (P.Q[i] * qmax[m] >= tau) idx.push_back(i);
                std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return P.Q[x] > P.Q[y]; });
            }
        });
        for (int m = 0; m < nmol; ++m) {
            for (int i : sig[m]) {
                P.s_sa.push_back(P.sa[i]);
                P.s_sb.push_back(P.sb[i]);
                P.s_poff.push_back(P.poff[i]);
                P.s_np.push_back(P.np[i]);
                P.s_Q.push_back(P.Q[i]);
                P.s_mol.push_back(m);
            }
            P.s_mbeg.push_back((int)P.s_sa.size());
        }
        SF_TRY(P.ds_sa.upload(P.s_sa));
        SF_TRY(P.ds_sb.upload(P.s_sb));
        SF_TRY(P.ds_poff.upload(P.s_poff));
        SF_TRY(P.ds_np.upload(P.s_np));
        SF_TRY(P.ds_Q.upload(P.s_Q));
        const long long n = (long long)P.s_sa.size();
        P.kmax = 0;
        for (int v : P.s_np) P.kmax = std::max(P.kmax, v);
        SF_TRY(P.d_kt.ensure((size_t)std::max(1LL, kKetFields * P.kmax * n) * sizeof(double)));
        SF_TRY(P.d_kc.ensure((size_t)std::max(1LL, 6 * n) * sizeof(double)));
        SF_TRY(launch_ket_transpose(P.sorted_dev(), P.kmax, b->d_prim.as<PrimPair>(),
                                    b->d_prim_q.as<double>(), b->d_prim_s.as<double>(),
                                    b->d_xyz.as<double>(), P.d_kt.as<double>(),
                                    P.d_kc.as<double>(), 0));
    }
    tm.mark("pair lists");
    // quartet lists per class (GPU builder), prefix sums on the host
    size_t maxbra = 1;
    for (auto& P : b->pc) maxbra = std::max(maxbra, P.s_sa.size());
    DevBuf d_nall, d_n64;
    SF_TRY(d_nall.ensure(maxbra * sizeof(int)));
    SF_TRY(d_n64.ensure(maxbra * sizeof(int)));
    SF_TRY(b->d_zero.upload(std::vector<int>(maxbra + 1, 0)));
    long long total_entries = 0;
    b->funcq = 0;
    b->primq = 0;
    for (int pcb = 0; pcb < kNumPairClass; ++pcb)
        for (int pck = 0; pck <= pcb; ++pck) {
            const int c = quartet_class(pcb, pck);
            QuartetClassData& Q = b->qc[c];
            const PairClassData& B = b->pc[pcb];
            const PairClassData& K = b->pc[pck];
            Q.pcb = pcb;
            Q.pck = pck;
            Q.same = (pcb == pck);
            Q.nbra = (int)B.s_sa.size();
            Q.ntile = class_ntile(c);
            Q.n_all.assign(Q.nbra, 0);
            Q.n64.assign(Q.nbra, 0);
            Q.ket0.assign(Q.nbra, 0);
            Q.nket.assign(Q.nbra, 0);
            for (int i = 0; i < Q.nbra; ++i) {
                const int m = B.s_mol[i];
                Q.ket0[i] = K.s_mbeg[m];
                Q.nket[i] = K.s_mbeg[m + 1] - K.s_mbeg[m];
            }
            SF_TRY(Q.d_ket0.upload(Q.ket0.empty() ? std::vector<int>(1, 0) : Q.ket0));
            SF_TRY(Q.d_nket.upload(Q.nket.empty() ? std::vector<int>(1, 0) : Q.nket));
            if (Q.nbra > 0 && !K.s_sa.empty()) {
                SF_TRY(launch_count_kets(B.ds_Q.as<double>(), Q.nbra, K.ds_Q.as<double>(),
                                         Q.d_ket0.as<int>(), Q.d_nket.as<int>(), Q.same, tau,
                                         tau_mixed, d_nall.as<int>(), d_n64.as<int>(), 0));
                SF_HIP(hipMemcpy(Q.n_all.data(), d_nall.p, Q.nbra * sizeof(int), hipMemcpyDeviceToHost));
                SF_HIP(hipMemcpy(Q.n64.data(), d_n64.p, Q.nbra * sizeof(int), hipMemcpyDeviceToHost));
            }
            Q.base_all.assign(Q.nbra + 1, 0);
            Q.base64.assign(Q.nbra + 1, 0);
            Q.base32.assign(Q.nbra + 1, 0);
            for (int i = 0; i < Q.nbra; ++i) {
                Q.base_all[i + 1] = Q.base_all[i] + Q.n_all[i];
                Q.base64[i + 1] = Q.base64[i] + Q.n64[i];
                Q.base32[i + 1] = Q.base32[i] + (Q.n_all[i] - Q.n64[i]);
            }
            Q.nq_all = Q.base_all[Q.nbra];
            Q.nq64 = Q.base64[Q.nbra];
            Q.nq32 = Q.base32[Q.nbra];
            SF_TRY(Q.d_base_all.upload(Q.base_all));
            SF_TRY(Q.d_base64.upload(Q.base64));
            SF_TRY(Q.d_base32.upload(Q.base32));
            SF_TRY(Q.d_k032.upload(Q.n64.empty() ? std::vector<int>(1, 0) : Q.n64));
            SF_TRY(Q.d_pos.ensure((size_t)std::max(1LL, Q.nq_all) * 6 * sizeof(unsigned)));
            if (class_split(c) || b->engine[c] == 1)
                SF_TRY(Q.d_raw.ensure((size_t)std::max(1LL, Q.nq_all) * class_nint(c) * sizeof(double)));
            else
                Q.d_raw.release();
            total_entries += 6 * Q.nq_all;
            // function-level unique quartets and primitive quartets, from
            // prefix sums over the ket list (a bra's kets are a contiguous run)
            const int la = pair_la(pcb), lb = pair_lb(pcb), lc = pair_la(pck), ld = pair_lb(pck);
            const long long nA = cart_n(la), nB = cart_n(lb), nC = cart_n(lc), nD = cart_n(ld);
            const size_t nk = K.s_sa.size();
            std::vector<long long> cum_f(nk + 1, 0), cum_p(nk + 1, 0);
            for (size_t j = 0; j < nk; ++j) {
                const bool eqcd = K.s_sa[j] == K.s_sb[j];
                cum_f[j + 1] = cum_f[j] + (eqcd ? nC * (nC + 1) / 2 : nC * nD);
                cum_p[j + 1] = cum_p[j] + K.s_np[j];
            }
            for (int i = 0; i < Q.nbra; ++i) {
                const bool eqab = B.s_sa[i] == B.s_sb[i];
                const long long pab = eqab ? nA * (nA + 1) / 2 : nA * nB;
                const int j0 = Q.ket0[i], j1 = Q.ket0[i] + Q.n_all[i];
                b->funcq += pab * (cum_f[j1] - cum_f[j0]);
                b->primq += (long long)B.s_np[i] * (cum_p[j1] - cum_p[j0]);
                // the diagonal quartet (i|i) has pab (pab + 1) / 2 unique function quartets
                if (Q.same && i >= j0 && i < j1) b->funcq -= pab * pab - pab * (pab + 1) / 2;
            }
        }
    tm.mark("quartet lists");
    // digestion plan: (destination, entry) pairs sorted by destination;
    // destinations are numbered per molecule (KeyMap)
    const int nmol_ = b->nmol;
    std::vector<unsigned long long> mkb(nmol_ + 1, 0);
    for (int m = 0; m < nmol_; ++m) {
        const unsigned long long ns = b->mol_shell[m + 1] - b->mol_shell[m];
        mkb[m + 1] = mkb[m] + 2 * ns * ns;
    }
    const unsigned long long nkeys = mkb[nmol_];
    KeyMap km;
    km.sh_mol = b->d_sh_mol.as<int>();
    km.sh_loc = b->d_sh_loc.as<int>();
    km.mol_ns = b->d_mol_ns.as<int>();
    km.mol_kbase = b->d_mol_kbase.as<unsigned>();
    const size_t E = (size_t)std::max(1LL, total_entries);
    DevBuf d_keys, d_vals, d_keys2, d_vals2, d_sort_tmp;
    SF_TRY(d_keys.ensure(E * sizeof(unsigned)));
    SF_TRY(d_vals.ensure(E * sizeof(unsigned long long)));
    SF_TRY(d_keys2.ensure(E * sizeof(unsigned)));
    SF_TRY(d_vals2.ensure(E * sizeof(unsigned long long)));
    long long ebase = 0;
    std::vector<ClassTab> tab(kNumQuartetClass);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        tab[c].pos = Q.d_pos.as<unsigned>();
        tab[c].qstride = Q.nq_all;
        if (Q.nq_all == 0) continue;
        SF_TRY(launch_plan_entries(c, b->pc[Q.pcb].sorted_dev(), b->pc[Q.pck].sorted_dev(),
                                   Q.d_ket0.as<int>(), Q.d_base_all.as<long long>(), Q.nq_all, km,
                                   d_keys.as<unsigned>() + ebase,
                                   d_vals.as<unsigned long long>() + ebase, 0));
        ebase += 6 * Q.nq_all;
    }
    DevBuf d_tab;
    SF_TRY(d_tab.upload(tab));
    SF_HIP(hipDeviceSynchronize());
    tm.mark("plan entries");
    const int nbits = key_bits(nkeys);
    size_t tmp_bytes = 0;
    SF_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_keys.as<unsigned>(),
                                     d_keys2.as<unsigned>(), d_vals.as<unsigned long long>(),
                                     d_vals2.as<unsigned long long>(), (size_t)total_entries, 0,
                                     nbits, (hipStream_t)0));
    SF_TRY(d_sort_tmp.ensure(tmp_bytes));
    if (total_entries > 0)
        SF_HIP(rocprim::radix_sort_pairs(d_sort_tmp.p, tmp_bytes, d_keys.as<unsigned>(),
                                         d_keys2.as<unsigned>(), d_vals.as<unsigned long long>(),
                                         d_vals2.as<unsigned long long>(), (size_t)total_entries,
                                         0, nbits, (hipStream_t)0));
    d_sort_tmp.release();
    d_keys.release();
    d_vals.release();
    SF_TRY(b->d_segb.ensure(nkeys * sizeof(long long)));
    SF_TRY(b->d_sege.ensure(nkeys * sizeof(long long)));
    SF_HIP(hipMemset(b->d_segb.p, 0, nkeys * sizeof(long long)));
    SF_HIP(hipMemset(b->d_sege.p, 0, nkeys * sizeof(long long)));
    SF_TRY(launch_segments(d_keys2.as<unsigned>(), total_entries, b->d_segb.as<long long>(),
                           b->d_sege.as<long long>(), 0));
    std::vector<long long> segb(nkeys), sege(nkeys);
    SF_HIP(hipMemcpy(segb.data(), b->d_segb.p, nkeys * sizeof(long long), hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(sege.data(), b->d_sege.p, nkeys * sizeof(long long), hipMemcpyDeviceToHost));
    tm.mark("sort + segments");
    // slot layout: destinations in key order, each a run of NE-double tiles.
    // Two passes over the molecules on host threads: counts per molecule and
    // destination group, then (after prefix sums) the lists in place.
    std::vector<unsigned long long> key_base(nkeys);
    std::vector<int> key_ne(nkeys);
    constexpr int NG = kNumTileGroup;
    std::vector<int> ds1[NG], ds2[NG], dxy[NG], ch_n[NG], ch_dest[NG], ch_part[NG];
    std::vector<unsigned long long> ch_src[NG];
    std::vector<int> m_dest[NG], m_pbase[NG], m_nch[NG];
    struct SlotCount {
        unsigned long long acc;
        long long nd[kNumTileGroup], nch[kNumTileGroup], nm[kNumTileGroup], npart[kNumTileGroup];
    };
    std::vector<SlotCount> cnt_m((size_t)nmol_ + 1, SlotCount{});
    auto key_shells = [&](int m, unsigned long long k, int& s1, int& s2, int& xy) {
        const int s0 = b->mol_shell[m];
        const unsigned long long ns = b->mol_shell[m + 1] - s0, n2 = ns * ns;
        const unsigned long long r = k - mkb[m];
        xy = (int)(r / n2);
        s1 = s0 + (int)((r % n2) / ns);
        s2 = s0 + (int)(r % ns);
    };
    parallel_ranges(nmol_, 4, [&](long long m0, long long m1) {
        for (long long m = m0; m < m1; ++m) {
            SlotCount c{};
            for (unsigned long long k = mkb[m]; k < mkb[m + 1]; ++k) {
                if (sege[k] <= segb[k]) continue;
                int s1, s2, xy;
                key_shells((int)m, k, s1, s2, xy);
                const long long n = sege[k] - segb[k];
                const int g = 4 * b->L[s1] + b->L[s2];
                const long long cw = reduce_chunk_tiles(cart_n(b->L[s1]) * cart_n(b->L[s2]));
                const long long nch = (n + cw - 1) / cw;
                c.acc += (unsigned long long)n * (cart_n(b->L[s1]) * cart_n(b->L[s2]));
                c.nd[g] += 1;
                c.nch[g] += nch;
                if (nch > 1) {
                    c.nm[g] += 1;
                    c.npart[g] += nch;
                }
            }
            cnt_m[m + 1] = c;
        }
    });
    // exclusive prefix sums: cnt_m[m] becomes molecule m's starting offsets
    for (int m = 0; m < nmol_; ++m) {
        SlotCount& n = cnt_m[m + 1];
        const SlotCount& p = cnt_m[m];
        n.acc += p.acc;
        for (int g = 0; g < NG; ++g) {
            n.nd[g] += p.nd[g];
            n.nch[g] += p.nch[g];
            n.nm[g] += p.nm[g];
            n.npart[g] += p.npart[g];
        }
    }
    const unsigned long long acc = cnt_m[nmol_].acc;
    if (acc >= (1ULL << 32)) {
        set_error(kEUnsupported, "digestion plan of %llu doubles exceeds 32-bit slots", acc);
        return kEUnsupported;
    }
    int npart[NG];
    for (int g = 0; g < NG; ++g) {
        const SlotCount& t = cnt_m[nmol_];
        if (t.nd[g] >= (1LL << 31) || t.nch[g] >= (1LL << 31) || t.npart[g] >= (1LL << 31)) {
            set_error(kEUnsupported, "digestion plan: too many destinations or chunks");
            return kEUnsupported;
        }
        npart[g] = (int)t.npart[g];
        ds1[g].resize(t.nd[g]);
        ds2[g].resize(t.nd[g]);
        dxy[g].resize(t.nd[g]);
        ch_src[g].resize(t.nch[g]);
        ch_n[g].resize(t.nch[g]);
        ch_dest[g].resize(t.nch[g]);
        ch_part[g].resize(t.nch[g]);
        m_dest[g].resize(t.nm[g]);
        m_pbase[g].resize(t.nm[g]);
        m_nch[g].resize(t.nm[g]);
    }
    parallel_ranges(nmol_, 4, [&](long long m0, long long m1) {
        for (long long m = m0; m < m1; ++m) {
            SlotCount o = cnt_m[m];
            for (unsigned long long k = mkb[m]; k < mkb[m + 1]; ++k) {
                int s1, s2, xy;
                key_shells((int)m, k, s1, s2, xy);
                key_ne[k] = cart_n(b->L[s1]) * cart_n(b->L[s2]);
                key_base[k] = o.acc;
                if (sege[k] <= segb[k]) continue;
                const long long n = sege[k] - segb[k];
                o.acc += (unsigned long long)n * key_ne[k];
                const int g = 4 * b->L[s1] + b->L[s2];
                const int d = (int)o.nd[g]++;
                ds1[g][d] = s1;
                ds2[g][d] = s2;
                dxy[g][d] = xy;
                // chunks of at most reduce_chunk_tiles(NE) tiles; several -> partial slots
                const long long cw = reduce_chunk_tiles(key_ne[k]);
                const int nch = (int)((n + cw - 1) / cw);
                const int pb = (int)o.npart[g];
                if (nch > 1) {
                    const long long j = o.nm[g]++;
                    m_dest[g][j] = d;
                    m_pbase[g][j] = pb;
                    m_nch[g][j] = nch;
                    o.npart[g] += nch;
                }
                for (int c = 0; c < nch; ++c) {
                    const long long j = o.nch[g]++;
                    ch_src[g][j] = key_base[k] + (unsigned long long)c * cw * key_ne[k];
                    ch_n[g][j] = (int)std::min<long long>(cw, n - (long long)c * cw);
                    ch_dest[g][j] = d;
                    ch_part[g][j] = nch > 1 ? pb + c : -1;
                }
            }
        }
    });
    tm.mark("slot layout + chunks");
    b->ntile_total = acc;
    SF_TRY(b->d_key_base.upload(key_base));
    DevBuf d_key_ne;
    SF_TRY(d_key_ne.upload(key_ne));
    SF_TRY(launch_plan_positions(d_keys2.as<unsigned>(), d_vals2.as<unsigned long long>(),
                                 total_entries, b->d_segb.as<long long>(),
                                 b->d_key_base.as<unsigned long long>(), d_key_ne.as<int>(),
                                 d_tab.as<ClassTab>(), 0));
    SF_HIP(hipDeviceSynchronize());
    d_keys2.release();
    d_vals2.release();
    SF_TRY(b->d_tiles.ensure(std::max<size_t>(1, acc) * sizeof(double)));
    long long part_total = 0;
    const int nf4[4] = {1, 3, 6, 10};
    for (int g = 0; g < NG; ++g) {
        b->ndest[g] = (int)ds1[g].size();
        b->nchunk[g] = (int)ch_n[g].size();
        b->nmulti[g] = (int)m_dest[g].size();
        b->part_off[g] = part_total;
        part_total += (long long)npart[g] * nf4[g / 4] * nf4[g % 4];
        SF_TRY(b->d_ds1[g].upload(ds1[g]));
        SF_TRY(b->d_ds2[g].upload(ds2[g]));
        SF_TRY(b->d_dxy[g].upload(dxy[g]));
        SF_TRY(b->d_ch_src[g].upload(ch_src[g]));
        SF_TRY(b->d_ch_n[g].upload(ch_n[g]));
        SF_TRY(b->d_ch_dest[g].upload(ch_dest[g]));
        SF_TRY(b->d_ch_part[g].upload(ch_part[g]));
        SF_TRY(b->d_m_dest[g].upload(m_dest[g]));
        SF_TRY(b->d_m_pbase[g].upload(m_pbase[g]));
        SF_TRY(b->d_m_nch[g].upload(m_nch[g]));
    }
    SF_TRY(b->d_part.ensure((size_t)std::max(1LL, part_total) * sizeof(double)));
    SF_HIP(hipDeviceSynchronize());
    tm.mark("positions + buffers");
    SF_TRY(build_far_lists(b));
    SF_HIP(hipDeviceSynchronize());
    tm.mark("near/far lists");
    b->screened = true;
    for (int c = 0; c < kNumQuartetClass; ++c) b->class_ms[c][0] = b->class_ms[c][1] = 0.0;
    return kOk;
}

extern "C" int sf_set_prim_tau(sf_basis* b, double tau_prim)
{
    if (!b || tau_prim != tau_prim) {
        set_error(kEInval, "sf_set_prim_tau: bad arguments");
        return kEInval;
    }
    b->tau_prim_set = tau_prim < 0.0 ? -1.0 : tau_prim;
    if (b->screened) b->tau_prim = b->tau_prim_set >= 0.0 ? b->tau_prim_set : 1e-8 * b->tau_mixed;
    b->drop_graph();
    return kOk;
}

extern "C" int sf_set_prim_tau_fp64(sf_basis* b, double tau_prim)
{
    if (!b || tau_prim != tau_prim) {
        set_error(kEInval, "sf_set_prim_tau_fp64: bad arguments");
        return kEInval;
    }
    b->tau_prim64_set = tau_prim < 0.0 ? -1.0 : tau_prim;
    if (b->screened) b->tau_prim64 = b->tau_prim64_set >= 0.0 ? b->tau_prim64_set : 1e-3 * b->tau;
    b->drop_graph();
    return kOk;
}

extern "C" int sf_quartet_counts(const sf_basis* b, long long* counts, long long* funcq,
                                 long long* primq)
{
    if (!b || !counts) {
        set_error(kEInval, "sf_quartet_counts: null argument");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_quartet_counts: call sf_screen first");
        return kEState;
    }
    for (int c = 0; c < kNumQuartetClass; ++c) {
        counts[3 * c] = b->qc[c].nq_all;
        counts[3 * c + 1] = b->qc[c].nq64;
        counts[3 * c + 2] = b->qc[c].nq32;
    }
    if (funcq) *funcq = b->funcq;
    if (primq) *primq = b->primq;
    return kOk;
}

extern "C" int sf_pair_list(const sf_basis* b, int pc, int* n, int* sa, int* sb, double* Q, int* nprim)
{
    if (!b || pc < 0 || pc >= kNumPairClass || !n) {
        set_error(kEInval, "sf_pair_list: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_pair_list: call sf_screen first");
        return kEState;
    }
    const PairClassData& P = b->pc[pc];
    *n = (int)P.s_sa.size();
    for (int i = 0; i < *n; ++i) {
        if (sa) sa[i] = P.s_sa[i];
        if (sb) sb[i] = P.s_sb[i];
        if (Q) Q[i] = P.s_Q[i];
        if (nprim) nprim[i] = P.s_np[i];
    }
    return kOk;
}

extern "C" int sf_quartet_list(const sf_basis* b, int c, int* nbra, int* n_all, int* n_fp64)
{
    if (!b || c < 0 || c >= kNumQuartetClass || !nbra) {
        set_error(kEInval, "sf_quartet_list: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_quartet_list: call sf_screen first");
        return kEState;
    }
    const QuartetClassData& Q = b->qc[c];
    *nbra = Q.nbra;
    for (int i = 0; i < Q.nbra; ++i) {
        if (n_all) n_all[i] = Q.n_all[i];
        if (n_fp64) n_fp64[i] = Q.n64[i];
    }
    return kOk;
}

// One class/segment launch with the class's engine.
// Tuning slot of segment seg of a build in `precision`.
static int tune_slot(int precision, int seg) { return precision == 0 ? 0 : 1 + seg; }

static int launch_class_engine(sf_basis* b, int c, int seg, int mode, const ClassParams& cp,
                               const Tune& tn, hipStream_t s)
{
    if (b->engine[c] == 1) return launch_eri_mmd(c, mode, cp, tn, s);
    const int fp32 = (seg == 1 && !tn.arith64) ? 1 : 0;
    QuartetClassData& Q = b->qc[c];
    if (mode != kModeJK || !Q.lists || cp.qlist)
        return launch_eri_class(c, fp32, mode, cp, tn, s);
    // one-node class in J/K mode: its near quartets, then its far ones (the
    // point-charge form), as two launches over the lists of cp's index space
    // (a prefix of cp.nq quartets, as the autotuner times, takes the same
    // share of each list)
    const int sp = cp.seg_base == Q.d_base_all.as<long long>() ? 0
                   : cp.seg_base == Q.d_base64.as<long long>() ? 1 : 2;
    const long long full = sp == 0 ? Q.nq_all : sp == 1 ? Q.nq64 : Q.nq32;
    // The far launch takes one thread per quartet in 256-thread blocks
    // whatever shape the autotuner gave the class: tuned together with the
    // near launch it inherited the near launch's shape, measured slower
    // (20-water FP64 build, same box: in order 6.95-6.97 -> 6.85-6.86 ms,
    // eager 5.96 -> 5.90 ms).
    Tune far_tune;
    far_tune.block = 256;
    far_tune.grid = 0;
    far_tune.arith64 = tn.arith64;
    for (int far = 0; far < 2; ++far) {
        ClassParams lc = cp;
        lc.qlist = Q.d_list[sp].as<unsigned>() + (far ? 0 : Q.nlist[sp][1]);
        lc.far = far;
        lc.nq = Q.nlist[sp][far];
        if (cp.nq < full) lc.nq = (long long)((double)lc.nq * (double)cp.nq / (double)full + 0.5);
        if (lc.nq > 0) SF_TRY(launch_eri_class(c, fp32, mode, lc, far ? far_tune : tn, s));
    }
    return kOk;
}

extern "C" int sf_set_streams(sf_basis* b, int nside)
{
    if (!b || nside < 0 || nside > sf_basis::kMaxSide) {
        set_error(kEInval, "sf_set_streams: nside must be 0..%d", sf_basis::kMaxSide);
        return kEInval;
    }
    b->nside = nside;
    b->nside_graph = nside;
    b->drop_graph();
    return kOk;
}

extern "C" int sf_set_engine(sf_basis* b, int qcls, int engine)
{
    if (!b || qcls < -1 || qcls >= kNumQuartetClass || engine < 0 || engine > 1) {
        set_error(kEInval, "sf_set_engine: bad arguments");
        return kEInval;
    }
    if (engine == 1 && qcls >= kNumClassSPD) {
        set_error(kEUnsupported, "sf_set_engine: quartet class %d has an f shell; the "
                  "McMurchie-Davidson engine covers s, p, d shells", qcls);
        return kEUnsupported;
    }
    for (int c = 0; c < kNumQuartetClass; ++c) {
        if (qcls >= 0 && c != qcls) continue;
        if (c >= kNumClassSPD) continue;  // f-shell classes: Rys only
        b->engine[c] = engine;
        QuartetClassData& Q = b->qc[c];
        if (b->screened && engine == 1)
            SF_TRY(Q.d_raw.ensure((size_t)std::max(1LL, Q.nq_all) * class_nint(c) * sizeof(double)));
    }
    b->drop_graph();
    return kOk;
}

static ClassParams class_params(sf_basis* b, int c, const double* Ps)
{
    QuartetClassData& Q = b->qc[c];
    ClassParams cp;
    cp.bra = b->pc[Q.pcb].sorted_dev();
    cp.ket = b->pc[Q.pck].sorted_dev();
    cp.prim = b->d_prim.as<PrimPair>();
    cp.shell_xyz = b->d_xyz.as<double>();
    cp.map = b->map();
    cp.seg_base = Q.d_base_all.as<long long>();
    cp.seg_k0 = nullptr;  // set per segment
    cp.ket0 = Q.d_ket0.as<int>();
    cp.qall_base = Q.d_base_all.as<long long>();
    cp.nq = Q.nq_all;
    cp.qstride = Q.nq_all;
    cp.same = Q.same;
    cp.P = Ps;
    cp.primq = b->d_prim_q.as<double>();
    cp.tau_prim = b->tau_prim64;  // set per segment
    cp.n64 = nullptr;
    cp.tau_prim_lo = 0.0;
    cp.out = nullptr;
    cp.raw = Q.d_raw.as<double>();
    cp.pos = Q.d_pos.as<unsigned>();
    cp.tiles = b->d_tiles.as<double>();
    cp.qlist = nullptr;
    cp.far = 0;
    cp.prims = b->d_prim_s.as<double>();
    const PairClassData& K = b->pc[Q.pck];
    cp.kt = K.d_kt.as<double>();
    cp.kt_n = (long long)K.s_sa.size();
    cp.kt_fs = (long long)K.kmax * cp.kt_n;
    cp.kc = K.d_kc.as<double>();
    return cp;
}

// The near and far lists of the one-node classes (QuartetClassData::d_list),
// built once per sf_screen from k_far_codes.
struct CodeIs {
    const unsigned char* code;
    unsigned char mask, want;
    __host__ __device__ bool operator()(unsigned qa) const { return (code[qa] & mask) == want; }
};

static int build_far_lists(sf_basis* b)
{
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        Q.lists = false;
        for (DevBuf& d : Q.d_list) d.release();
        if (class_ltot(c) > 1 || Q.nq_all == 0) continue;
        if (Q.nq_all >= (1LL << 32)) continue;  // 32-bit list entries
        const ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        DevBuf codes, d_cnt, tmp;
        SF_TRY(codes.ensure((size_t)Q.nq_all));
        SF_TRY(d_cnt.ensure(sizeof(unsigned long long)));
        SF_TRY(launch_far_codes(cp, Q.nq_all, Q.d_k032.as<int>(), codes.as<unsigned char>(), 0));
        // index space sp: mask and wanted bit 1 (FP64 segment); far: bit 0
        const unsigned char mask[3] = {1, 3, 3}, seg[3] = {0, 2, 0};
        const long long nsp[3] = {Q.nq_all, Q.nq64, Q.nq32};
        for (int sp = 0; sp < 3; ++sp) {
            SF_TRY(Q.d_list[sp].ensure((size_t)std::max(1LL, nsp[sp]) * sizeof(unsigned)));
            long long at = 0;
            for (int far = 1; far >= 0; --far) {
                const CodeIs pred{codes.as<unsigned char>(), mask[sp], (unsigned char)(seg[sp] | far)};
                rocprim::counting_iterator<unsigned> in(0u);
                unsigned* out = Q.d_list[sp].as<unsigned>() + at;
                size_t bytes = 0;
                SF_HIP(rocprim::select(nullptr, bytes, in, out, d_cnt.as<unsigned long long>(),
                                       (size_t)Q.nq_all, pred, (hipStream_t)0));
                SF_TRY(tmp.ensure(bytes));
                SF_HIP(rocprim::select(tmp.p, bytes, in, out, d_cnt.as<unsigned long long>(),
                                       (size_t)Q.nq_all, pred, (hipStream_t)0));
                unsigned long long n = 0;
                SF_HIP(hipMemcpy(&n, d_cnt.p, sizeof(n), hipMemcpyDeviceToHost));
                Q.nlist[sp][far] = (long long)n;
                at += (long long)n;
            }
            if (at != nsp[sp]) {
                set_error(kEState, "near/far lists of class %d: %lld of %lld quartets", c, at,
                          nsp[sp]);
                return kEState;
            }
        }
        Q.lists = true;
    }
    return kOk;
}

// Segment `seg` of quartet class c: 0 = all quartets (precision 0) or the
// FP64 part (precision 1), 1 = the FP32 part.  Zero-ket offsets come from a
// per-class zero array (d_k032 holds n64, the FP32 segment's first ket).
static void set_segment(sf_basis* b, int c, int precision, int seg, ClassParams& cp,
                        const int* zeros)
{
    QuartetClassData& Q = b->qc[c];
    cp.tau_prim = b->tau_prim64;
    if (precision == 0) {
        cp.seg_base = Q.d_base_all.as<long long>();
        cp.seg_k0 = zeros;
        cp.nq = Q.nq_all;
    } else if (seg == 0) {
        cp.seg_base = Q.d_base64.as<long long>();
        cp.seg_k0 = zeros;
        cp.nq = Q.nq64;
    } else {
        cp.seg_base = Q.d_base32.as<long long>();
        cp.seg_k0 = Q.d_k032.as<int>();
        cp.nq = Q.nq32;
        cp.tau_prim = b->tau_prim;
    }
}

extern "C" int sf_prim_quartet_counts(sf_basis* b, int precision, long long* counts)
{
    if (!b || !counts || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_prim_quartet_counts: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_prim_quartet_counts: call sf_screen first");
        return kEState;
    }
    const int nblock = 2048;
    DevBuf d_part;
    SF_TRY(d_part.ensure(nblock * sizeof(unsigned long long)));
    std::vector<unsigned long long> part(nblock);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        ClassParams cp = class_params(b, c, nullptr);
        for (int seg = 0; seg < 2; ++seg) {
            counts[2 * c + seg] = 0;
            if (seg > precision || Q.nq_all == 0) continue;
            set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
            if (cp.nq == 0) continue;
            SF_TRY(launch_count_prims(cp, nblock, d_part.as<unsigned long long>(), 0));
            SF_HIP(hipMemcpy(part.data(), d_part.p, nblock * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            unsigned long long t = 0;
            for (unsigned long long v : part) t += v;
            counts[2 * c + seg] = (long long)t;
        }
    }
    return kOk;
}

extern "C" int sf_far_prim_quartet_counts(sf_basis* b, int precision, long long* counts)
{
    if (!b || !counts || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_far_prim_quartet_counts: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_far_prim_quartet_counts: call sf_screen first");
        return kEState;
    }
    const int nblock = 2048;
    DevBuf d_part;
    SF_TRY(d_part.ensure(nblock * sizeof(unsigned long long)));
    std::vector<unsigned long long> part(nblock);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        ClassParams cp = class_params(b, c, nullptr);
        for (int seg = 0; seg < 2; ++seg) {
            counts[2 * c + seg] = 0;
            if (seg > precision || !Q.lists) continue;
            set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
            // the far list of the segment's index space
            const int sp = precision == 0 ? 0 : 1 + seg;
            cp.qlist = Q.d_list[sp].as<unsigned>();
            cp.nq = Q.nlist[sp][1];
            if (cp.nq == 0) continue;
            SF_TRY(launch_count_prims(cp, nblock, d_part.as<unsigned long long>(), 0));
            SF_HIP(hipMemcpy(part.data(), d_part.p, nblock * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            unsigned long long t = 0;
            for (unsigned long long v : part) t += v;
            counts[2 * c + seg] = (long long)t;
        }
    }
    return kOk;
}

extern "C" int sf_eri_dense(sf_basis* b, double* eri)
{
    if (!b || !eri) {
        set_error(kEInval, "sf_eri_dense: null argument");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_eri_dense: call sf_screen first");
        return kEState;
    }
    if (b->nmol != 1) {
        set_error(kEUnsupported, "sf_eri_dense: one molecule only");
        return kEUnsupported;
    }
    const long long n = b->nbf;
    if (n > 256) {
        set_error(kEUnsupported, "sf_eri_dense: N = %lld is too large for a dense tensor", n);
        return kEUnsupported;
    }
    std::fill(eri, eri + n * n * n * n, 0.0);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        if (Q.nq_all == 0) continue;
        const int nint = class_nint(c);
        DevBuf d_dense;
        SF_TRY(d_dense.ensure((size_t)Q.nq_all * nint * sizeof(double)));
        ClassParams cp = class_params(b, c, nullptr);
        set_segment(b, c, 0, 0, cp, b->d_zero.as<int>());
        cp.out = d_dense.as<double>();
        SF_TRY(launch_class_engine(b, c, 0, kModeRaw, cp, Tune(), 0));
        std::vector<double> raw((size_t)Q.nq_all * nint);
        SF_HIP(hipMemcpy(raw.data(), d_dense.p, raw.size() * sizeof(double), hipMemcpyDeviceToHost));
        const PairClassData& B = b->pc[Q.pcb];
        const PairClassData& K = b->pc[Q.pck];
        const int nA = cart_n(pair_la(Q.pcb)), nB = cart_n(pair_lb(Q.pcb));
        const int nC = cart_n(pair_la(Q.pck)), nD = cart_n(pair_lb(Q.pck));
        for (int i = 0; i < Q.nbra; ++i)
            for (int j = 0; j < Q.n_all[i]; ++j) {
                const long long q = Q.base_all[i] + j;
                const int fa = b->foff[B.s_sa[i]], fb = b->foff[B.s_sb[i]];
                const int fc = b->foff[K.s_sa[j]], fd = b->foff[K.s_sb[j]];
                int k = 0;
                for (int ia = 0; ia < nA; ++ia)
                    for (int ib = 0; ib < nB; ++ib)
                        for (int ic = 0; ic < nC; ++ic)
                            for (int id = 0; id < nD; ++id, ++k) {
                                const long long I = fa + ia, J = fb + ib, Kk = fc + ic, Ll = fd + id;
                                const double v = raw[(size_t)k * Q.nq_all + q] * b->fnorm[I] *
                                                 b->fnorm[J] * b->fnorm[Kk] * b->fnorm[Ll];
                                const long long P8[8][4] = {{I, J, Kk, Ll}, {J, I, Kk, Ll},
                                                            {I, J, Ll, Kk}, {J, I, Ll, Kk},
                                                            {Kk, Ll, I, J}, {Ll, Kk, I, J},
                                                            {Kk, Ll, J, I}, {Ll, Kk, J, I}};
                                for (auto& p : P8) eri[((p[0] * n + p[1]) * n + p[2]) * n + p[3]] = v;
                            }
            }
    }
    return kOk;
}

static int check_jk_args(sf_basis* b, const double* P, double* J, double* K, int precision)
{
    if (!b || !P || !J || !K || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_jk: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_jk: call sf_screen first");
        return kEState;
    }
    return kOk;
}

// J/K driver (kernel launches only: no allocation, no synchronisation, no
// memset, so a stream capture of it is a valid graph); when ms is non-null,
// HIP events bracket every launch group on the stream and ms receives
// milliseconds per group: [2c] FP64 and [2c+1] FP32 ERI+digestion kernels
// of quartet class c (c = 0..54), [kSlotScale] density scaling,
// [kSlotReduce] plan reduction + finalisation.
static int run_jk(sf_basis* b, const double* P, double* J, double* K, int precision,
                  hipStream_t s, double* ms, bool capture = false)
{
    SF_TRY(check_jk_args(b, P, J, K, precision));
    const MolLayout ml = b->layout();
    std::vector<hipEvent_t> ev;
    auto mark = [&]() -> int {
        if (!ms) return kOk;
        hipEvent_t e;
        SF_HIP(hipEventCreate(&e));
        SF_HIP(hipEventRecord(e, s));
        ev.push_back(e);
        return kOk;
    };
    std::vector<int> slot;  // ms index of the interval ending at each mark
    SF_TRY(mark());
    SF_TRY(launch_scale_density(P, b->d_fnorm.as<double>(), ml, b->d_Ps.as<double>(), s));
    SF_TRY(mark());
    slot.push_back(kSlotScale);
    // concurrent launches (not when timing launch groups): the classes write
    // disjoint tile slots and raw planes, the reductions disjoint blocks of
    // X and Y, so each group can go to any stream between a fork and a join
    const int nside = ms ? 0 : capture ? b->nside_graph : b->nside;
    if (nside > 0) SF_TRY(b->ensure_side());
    hipStream_t st[1 + sf_basis::kMaxSide] = {s};
    for (int k = 0; k < nside; ++k) st[1 + k] = b->side[k];
    auto fork = [&]() -> int {
        if (nside == 0) return kOk;
        SF_HIP(hipEventRecord(b->ev_fork, s));
        for (int k = 0; k < nside; ++k) SF_HIP(hipStreamWaitEvent(b->side[k], b->ev_fork, 0));
        return kOk;
    };
    auto join = [&]() -> int {
        for (int k = 0; k < nside; ++k) {
            SF_HIP(hipEventRecord(b->ev_join[k], b->side[k]));
            SF_HIP(hipStreamWaitEvent(s, b->ev_join[k], 0));
        }
        return kOk;
    };
    // longest-first onto the least-loaded stream, by a cost estimate
    double load[1 + sf_basis::kMaxSide] = {0.0};
    auto pick = [&](double cost) -> hipStream_t {
        int k = 0;
        for (int t = 1; t <= nside; ++t)
            if (load[t] < load[k]) k = t;
        load[k] += cost;
        return st[k];
    };
    int order[kNumQuartetClass];
    for (int c = 0; c < kNumQuartetClass; ++c) order[c] = c;
    bool measured = true;
    for (int c = 0; c < kNumQuartetClass; ++c)
        if (b->qc[c].nq_all > 0 && !(b->class_ms[c][precision] > 0.0)) measured = false;
    auto class_cost = [&](int c) {
        return measured ? b->class_ms[c][precision]
                        : (double)b->qc[c].nq_all * (double)class_nint(c) + 2.0e5;
    };
    if (nside > 0)
        std::stable_sort(order, order + kNumQuartetClass,
                         [&](int x, int y) { return class_cost(x) > class_cost(y); });
    auto one_class = [&](int c, hipStream_t cs) -> int {
        QuartetClassData& Q = b->qc[c];
        ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        if (precision == 1 && b->engine[c] == 0 &&
            (b->tune[c][2].arith64 || class_generic(c))) {
            // both segments in FP64 arithmetic: one launch over every quartet
            // of the class, each bra's FP32-segment kets with that segment's
            // primitive threshold (the same integrals as two launches)
            set_segment(b, c, 0, 0, cp, b->d_zero.as<int>());
            cp.tau_prim = b->tau_prim64;
            cp.n64 = Q.d_k032.as<int>();
            cp.tau_prim_lo = b->tau_prim;
            SF_TRY(launch_class_engine(b, c, 0, kModeJK, cp, b->tune[c][1], cs));
            SF_TRY(mark());
            slot.push_back(2 * c);
            return kOk;
        }
        if (precision == 1 && b->engine[c] == 0 && class_split(c)) {
            // split class, mixed build: the two segments write their raw
            // planes at class-wide quartet indices, then one digestion
            // launch over the class (as in an FP64 build) reads them all
            for (int seg = 0; seg <= 1; ++seg) {
                set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
                SF_TRY(launch_class_engine(b, c, seg, kModeEriOnly, cp, b->tune[c][1 + seg], cs));
                if (seg == 0) {
                    SF_TRY(mark());
                    slot.push_back(2 * c);
                }
            }
            set_segment(b, c, 0, 0, cp, b->d_zero.as<int>());
            SF_TRY(launch_class_engine(b, c, 0, kModeDigestOnly, cp, b->tune[c][0], cs));
            SF_TRY(mark());
            slot.push_back(2 * c + 1);
            return kOk;
        }
        for (int seg = 0; seg <= precision; ++seg) {
            set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
            SF_TRY(launch_class_engine(b, c, seg, kModeJK, cp, b->tune[c][tune_slot(precision, seg)], cs));
            SF_TRY(mark());
            slot.push_back(2 * c + seg);
        }
        return kOk;
    };
    SF_TRY(fork());
    int rc = kOk;
    for (int oc = 0; oc < kNumQuartetClass && rc == kOk; ++oc) {
        const int c = order[oc];
        if (b->qc[c].nq_all == 0) continue;
        rc = one_class(c, nside ? pick(class_cost(c)) : s);
    }
    SF_TRY(join());
    SF_TRY(rc);
    SF_TRY(launch_fill(b->d_X.as<double>(), b->extent, 0.0, s));
    SF_TRY(launch_fill(b->d_Y.as<double>(), b->extent, 0.0, s));
    SF_TRY(fork());
    const int nf[4] = {1, 3, 6, 10};
    for (int k = 0; k <= nside; ++k) load[k] = 0.0;
    for (int g = 0; g < kNumTileGroup && rc == kOk; ++g) {
        DestList dl;
        dl.s1 = b->d_ds1[g].as<int>();
        dl.s2 = b->d_ds2[g].as<int>();
        dl.xy = b->d_dxy[g].as<int>();
        dl.ch_src = b->d_ch_src[g].as<unsigned long long>();
        dl.ch_n = b->d_ch_n[g].as<int>();
        dl.ch_dest = b->d_ch_dest[g].as<int>();
        dl.ch_part = b->d_ch_part[g].as<int>();
        dl.nch = b->nchunk[g];
        dl.m_dest = b->d_m_dest[g].as<int>();
        dl.m_pbase = b->d_m_pbase[g].as<int>();
        dl.m_nch = b->d_m_nch[g].as<int>();
        dl.nm = b->nmulti[g];
        dl.part = b->d_part.as<double>() + b->part_off[g];
        rc = launch_reduce(nf[g / 4], nf[g % 4], dl, b->d_tiles.as<double>(), b->map(),
                           b->d_X.as<double>(), b->d_Y.as<double>(),
                           nside ? pick((double)b->nchunk[g] * nf[g / 4] * nf[g % 4]) : s);
    }
    SF_TRY(join());
    SF_TRY(rc);
    SF_TRY(launch_finalize(b->d_X.as<double>(), b->d_Y.as<double>(), b->d_fnorm.as<double>(), ml, J,
                           K, s));
    SF_TRY(mark());
    slot.push_back(kSlotReduce);
    if (ms) {
        for (int k = 0; k < kNumSlots; ++k) ms[k] = 0.0;
        SF_HIP(hipEventSynchronize(ev.back()));
        for (size_t k = 1; k < ev.size(); ++k) {
            float t = 0.f;
            SF_HIP(hipEventElapsedTime(&t, ev[k - 1], ev[k]));
            ms[slot[k - 1]] += t;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (int c = 0; c < kNumQuartetClass; ++c)
            b->class_ms[c][precision] = b->qc[c].nq_all > 0 ? ms[2 * c] + ms[2 * c + 1] : 0.0;
    }
    return kOk;
}

extern "C" int sf_jk(sf_basis* b, const double* P, double* J, double* K, int precision,
                     void* stream)
{
    return run_jk(b, P, J, K, precision, (hipStream_t)stream, nullptr);
}

extern "C" int sf_jk_profile(sf_basis* b, const double* P, double* J, double* K, int precision,
                             void* stream, double* ms)
{
    if (!ms) {
        set_error(kEInval, "sf_jk_profile: null ms");
        return kEInval;
    }
    return run_jk(b, P, J, K, precision, (hipStream_t)stream, ms);
}

// ------------------------------------------------------------ graphs ----
// The whole J/K build (density scaling, every class kernel, plan
// reduction, finalisation) captured once into a hipGraph and replayed.
// The capture runs on a private stream (capture needs a non-default
// stream); the replay goes on the caller's stream.  A new capture is made
// when P/J/K/precision change and after sf_screen, sf_autotune,
// sf_set_tuning or sf_basis_set_layout.
extern "C" int sf_jk_graph(sf_basis* b, const double* P, double* J, double* K, int precision,
                           void* stream)
{
    SF_TRY(check_jk_args(b, P, J, K, precision));
    JKGraph& g = b->graph;
    if (!g.exec || g.P != P || g.J != J || g.K != K || g.precision != precision) {
        b->drop_graph();
        if (!b->cap_stream) SF_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        SF_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = run_jk(b, P, J, K, precision, b->cap_stream, nullptr, true);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap_stream, &graph);
        if (rc != kOk) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        SF_HIP(e);
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        SF_HIP(ei);
        g.exec = exec;
        g.P = P;
        g.J = J;
        g.K = K;
        g.precision = precision;
    }
    SF_HIP(hipGraphLaunch(g.exec, (hipStream_t)stream));
    return kOk;
}

// ---------------------------------------------------------- autotuner ----
// Launch shapes tried per quartet class and precision segment: 64/128/256
// threads per block, each with one thread (or wavefront) per work item
// (grid 0) or a persistent grid of 2 or 8 blocks per compute unit.  The
// class kernel runs once untimed (caches, code load), then each candidate
// once between HIP events on the stream; the fastest wins.
static constexpr long long kTuneQuartets = 1LL << 20;

extern "C" int sf_autotune(sf_basis* b, const double* P, int precision, void* stream, int* shapes,
                           double* ms)
{
    if (!b || !P || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_autotune: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_autotune: call sf_screen first");
        return kEState;
    }
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 0;
    SF_HIP(hipGetDevice(&dev));
    SF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu <= 0) ncu = 1;
    b->drop_graph();
    SF_TRY(launch_scale_density(P, b->d_fnorm.as<double>(), b->layout(), b->d_Ps.as<double>(), s));
    hipEvent_t e0, e1;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    int rc = kOk;
    const int blocks[3] = {64, 128, 256};
    const int mults[3] = {0, 2, 8};
    // fastest launch shape for launches of (segment seg of) class c with
    // parameters tcp in arithmetic ar; candidates are timed on a prefix of
    // the quartets: the ranking of launch shapes does not need every quartet,
    // and this bounds the tuning time for the largest classes
    auto time_shapes = [&](int c, int seg, ClassParams tcp, int ar, Tune& best, float& best_ms) -> int {
        tcp.nq = std::min(tcp.nq, kTuneQuartets);
        // work items of one launch: quartets, or waves x component slices
        const long long items =
            class_split(c) ? ((tcp.nq + 63) / 64) * 64 * class_nslice(c) : tcp.nq;
        best_ms = 1e30f;
        Tune warm;
        warm.arith64 = ar;
        SF_TRY(launch_class_engine(b, c, seg, kModeJK, tcp, warm, s));  // untimed warm-up
        for (int bi = 0; bi < 3; ++bi)
            for (int mi = 0; mi < 3; ++mi) {
                // the run-time-L classes keep 256-thread blocks (timed once,
                // for the stream balance of run_jk)
                if (class_generic(c) && (bi != 2 || mi != 0)) continue;
                Tune t;
                t.block = blocks[bi];
                t.grid = mults[mi] * ncu;
                t.arith64 = ar;
                const long long need = (items + t.block - 1) / t.block;
                if (t.grid > 0 && t.grid >= need) continue;  // same as grid 0
                if (hipEventRecord(e0, s) != hipSuccess) {
                    set_error(kEHip, "sf_autotune: event record failed");
                    return kEHip;
                }
                SF_TRY(launch_class_engine(b, c, seg, kModeJK, tcp, t, s));
                float el = 0.f;
                if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&el, e0, e1) != hipSuccess) {
                    set_error(kEHip, "sf_autotune: event timing failed");
                    return kEHip;
                }
                if (el < best_ms) {
                    best_ms = el;
                    best = t;
                }
            }
        return kOk;
    };
    // time of a whole launch over n quartets from its prefix timing
    auto full = [](float t, long long n) {
        return (double)t * (double)n / (double)std::max(1LL, std::min(n, kTuneQuartets));
    };
    for (int c = 0; c < kNumQuartetClass && rc == kOk; ++c) {
        QuartetClassData& Q = b->qc[c];
        ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        float seg_ms[2] = {0.f, 0.f};
        float ms32 = 0.f;  // FP32 segment in FP32 arithmetic
        Tune best32;
        for (int seg = 0; seg <= precision && rc == kOk; ++seg) {
            set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
            const int slot = tune_slot(precision, seg);
            if (ms) ms[3 * c + slot] = 0.0;
            if (Q.nq_all == 0 || cp.nq == 0) continue;
            Tune best = b->tune[c][slot];
            float best_ms = 1e30f;
            rc = time_shapes(c, seg, cp, 0, best, best_ms);
            ms32 = best_ms;
            best32 = best;
            // the FP32 segment also tries FP64 arithmetic (never less accurate)
            if (!rc && seg == 1 && b->engine[c] == 0) {
                Tune b64;
                float m64 = 1e30f;
                rc = time_shapes(c, seg, cp, 1, b64, m64);
                if (!rc && m64 < best_ms) {
                    best_ms = m64;
                    best = b64;
                }
            }
            if (rc) break;
            b->tune[c][slot] = best;
            seg_ms[seg] = best_ms;
            if (ms) ms[3 * c + slot] = best_ms;
        }
        if (!rc && Q.nq_all > 0)
            b->class_ms[c][precision] = precision == 0 ? full(seg_ms[0], Q.nq_all)
                                                       : full(seg_ms[0], Q.nq64) + full(seg_ms[1], Q.nq32);
        // FP64 arithmetic in the FP32 segment means one launch over both
        // segments (run_jk), so the choice is between that launch and two
        // with FP32 arithmetic
        if (!rc && precision == 1 && b->engine[c] == 0 && Q.nq64 > 0 && Q.nq32 > 0) {
            ClassParams mcp = class_params(b, c, b->d_Ps.as<double>());
            set_segment(b, c, 0, 0, mcp, b->d_zero.as<int>());
            mcp.tau_prim = b->tau_prim64;
            mcp.n64 = Q.d_k032.as<int>();
            mcp.tau_prim_lo = b->tau_prim;
            Tune bm;
            float mm = 1e30f;
            rc = time_shapes(c, 0, mcp, 0, bm, mm);
            if (rc) break;
            const double two = full(seg_ms[0], Q.nq64) + full(ms32, Q.nq32);
            if (full(mm, Q.nq_all) < two) {
                b->tune[c][1] = bm;
                b->tune[c][2].arith64 = 1;
                if (ms) ms[3 * c + 1] = mm, ms[3 * c + 2] = 0.0;
                b->class_ms[c][1] = full(mm, Q.nq_all);
            } else {
                b->tune[c][2] = best32;
                if (ms) ms[3 * c + 2] = ms32;
                b->class_ms[c][1] = two;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (shapes)
        for (int c = 0; c < kNumQuartetClass; ++c)
            for (int k = 0; k < 3; ++k) {
                shapes[9 * c + 3 * k] = b->tune[c][k].block;
                shapes[9 * c + 3 * k + 1] = b->tune[c][k].grid;
                shapes[9 * c + 3 * k + 2] = b->tune[c][k].arith64;
            }
    return kOk;
}

extern "C" int sf_set_tuning(sf_basis* b, const int* shapes)
{
    if (!b || !shapes) {
        set_error(kEInval, "sf_set_tuning: null argument");
        return kEInval;
    }
    for (int k = 0; k < 3 * kNumQuartetClass; ++k) {
        const int block = shapes[3 * k], grid = shapes[3 * k + 1], ar = shapes[3 * k + 2];
        if (block < 64 || block > 256 || block % 64 != 0 || grid < 0 || ar < 0 || ar > 1) {
            set_error(kEInval, "sf_set_tuning: entry %d: block %d grid %d arith %d", k, block, grid, ar);
            return kEInval;
        }
    }
    for (int c = 0; c < kNumQuartetClass; ++c)
        for (int k = 0; k < 3; ++k) {
            b->tune[c][k].block = shapes[9 * c + 3 * k];
            b->tune[c][k].grid = shapes[9 * c + 3 * k + 1];
            b->tune[c][k].arith64 = shapes[9 * c + 3 * k + 2];
        }
    b->drop_graph();
    return kOk;
}

extern "C" int sf_basis_set_nuclei(sf_basis* b, int natom, const double* Z, const double* xyz)
{
    if (!b || natom < 0 || (natom > 0 && (!Z || !xyz))) {
        set_error(kEInval, "sf_basis_set_nuclei: bad arguments");
        return kEInval;
    }
    for (int a = 0; a < natom; ++a) {
        if (!(Z[a] >= 0.0) || !std::isfinite(Z[a]) || !std::isfinite(xyz[3 * a]) ||
            !std::isfinite(xyz[3 * a + 1]) || !std::isfinite(xyz[3 * a + 2])) {
            set_error(kEInval, "sf_basis_set_nuclei: nucleus %d has charge %g at (%g, %g, %g)", a,
                      Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
            return kEInval;
        }
    }
    b->nnuc = natom;
    SF_TRY(b->d_Z.upload(std::vector<double>(Z, Z + natom)));
    SF_TRY(b->d_nuc.upload(std::vector<double>(xyz, xyz + 3 * natom)));
    return kOk;
}

extern "C" int sf_one_electron(sf_basis* b, double* S, double* T, double* V, void* stream)
{
    if (!b || !S || !T || !V) {
        set_error(kEInval, "sf_one_electron: null argument");
        return kEInval;
    }
    if (b->nnuc < b->mol_atom[b->nmol]) {
        set_error(kEState, "sf_one_electron: %d nuclei set, the molecules need %d "
                  "(sf_basis_set_nuclei)", b->nnuc, b->mol_atom[b->nmol]);
        return kEState;
    }
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        SF_TRY(launch_one_electron(p, P.d_sa.as<int>(), P.d_sb.as<int>(), P.d_poff.as<int>(),
                                   P.d_np.as<int>(), (int)P.sa.size(), b->d_prim.as<PrimPair>(),
                                   b->d_xyz.as<double>(), b->d_foff.as<int>(),
                                   b->d_fnorm.as<double>(), b->map(), b->d_sh_mol.as<int>(),
                                   b->d_mol_nuc.as<int>(), b->d_Z.as<double>(),
                                   b->d_nuc.as<double>(), S, T, V, (hipStream_t)stream));
    }
    return kOk;
}

// Layout of the gradient partials and the per-atom lists that sum them.
static int prepare_grad(sf_basis* b)
{
    if (b->grad_ready) return kOk;
    const int natom = b->natom;
    b->gpart_bra.assign(kNumQuartetClass, -1);
    b->gpart_ket.assign(kNumQuartetClass, -1);
    long long off = 0;
    size_t gout = 1;
    std::vector<std::vector<long long>> lists(natom);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        const QuartetClassData& Q = b->qc[c];
        if (Q.nq_all == 0) continue;
        const PairClassData& B = b->pc[Q.pcb];
        const PairClassData& K = b->pc[Q.pck];
        b->gpart_bra[c] = off;
        for (int i = 0; i < Q.nbra; ++i) {
            lists[b->atom[B.s_sa[i]]].push_back(off + 6LL * i);
            lists[b->atom[B.s_sb[i]]].push_back(off + 6LL * i + 3);
        }
        off += 6LL * Q.nbra;
        b->gpart_ket[c] = off;
        const int nket = (int)K.s_sa.size();
        for (int j = 0; j < nket; ++j) {
            lists[b->atom[K.s_sa[j]]].push_back(off + 6LL * j);
            lists[b->atom[K.s_sb[j]]].push_back(off + 6LL * j + 3);
        }
        off += 6LL * nket;
        gout = std::max(gout, (size_t)class_nslice(c) * 9 * (size_t)Q.nq_all);
    }
    std::vector<long long> csr_off(natom + 1, 0), csr_idx;
    for (int a = 0; a < natom; ++a) {
        csr_off[a + 1] = csr_off[a] + (long long)lists[a].size();
        csr_idx.insert(csr_idx.end(), lists[a].begin(), lists[a].end());
    }
    SF_TRY(b->d_gpart.ensure(std::max<long long>(1, off) * sizeof(double)));
    SF_TRY(b->d_gout.ensure(gout * sizeof(double)));
    SF_TRY(b->d_gcsr_off.upload(csr_off));
    SF_TRY(b->d_gcsr_idx.upload(csr_idx.empty() ? std::vector<long long>(1, 0) : csr_idx));
    long long npair = 0;
    for (auto& P : b->pc) npair += (long long)P.sa.size();
    SF_TRY(b->d_grows.ensure(std::max<long long>(1, npair) * 3 * std::max(1, natom) * sizeof(double)));
    SF_TRY(b->d_gonee.ensure(std::max<long long>(1, npair) * 6 * std::max(1, natom) * sizeof(double)));
    long long maxq = 1;
    for (int c = 0; c < kNumQuartetClass; ++c) maxq = std::max(maxq, b->qc[c].nq_all);
    SF_TRY(b->d_gkc.ensure((size_t)maxq * 12 * sizeof(double)));
    b->gkc_q = maxq;
    SF_TRY(b->d_Ws.ensure((size_t)std::max(1LL, b->extent) * sizeof(double)));
    SF_TRY(b->d_g2.ensure(3 * std::max(1, natom) * sizeof(double)));
    SF_TRY(b->d_gtot.ensure(3 * std::max(1, natom) * sizeof(double)));
    b->grad_ready = true;
    return kOk;
}

extern "C" int sf_grad(sf_basis* b, const double* P, const double* W, double* grad, int precision,
                       void* stream)
{
    if (!b || !P || !W || !grad || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_grad: null argument");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_grad: call sf_screen first");
        return kEState;
    }
    if (std::find(b->L.begin(), b->L.end(), 3) != b->L.end()) {
        set_error(kEUnsupported, "sf_grad: nuclear gradients cover s, p, d shells; the basis has "
                  "f shells");
        return kEUnsupported;
    }
    if (b->nmol != 1) {
        set_error(kEUnsupported, "sf_grad: one molecule only");
        return kEUnsupported;
    }
    if (b->nnuc != b->natom || b->natom <= 0) {
        set_error(kEState, "sf_grad: nuclei (sf_basis_set_nuclei) must match the %d atoms", b->natom);
        return kEState;
    }
    if (!b->have_shell_atom) {
        set_error(kEState, "sf_grad: the basis was created without shell_atom (no shell-to-atom map)");
        return kEState;
    }
    hipStream_t s = (hipStream_t)stream;
    SF_TRY(prepare_grad(b));
    const int natom = b->natom;
    const MolLayout ml = b->layout();
    SF_TRY(launch_scale_density(P, b->d_fnorm.as<double>(), ml, b->d_Ps.as<double>(), s));
    SF_TRY(launch_scale_density(W, b->d_fnorm.as<double>(), ml, b->d_Ws.as<double>(), s));
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        if (Q.nq_all == 0) continue;
        ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        // every segment writes its quartets' partials at their class-wide
        // index, so the reductions below see one array either way.  Mixed
        // gradients have two ways to evaluate a class with FP32 quartets:
        // its two segments as two launches (the FP32 one in FP32
        // arithmetic), or one launch over the class in FP64 arithmetic
        // with each quartet's own segment threshold (never less accurate;
        // on this GPU, without packed FP32, it can be the faster one).  The
        // first mixed sf_grad after sf_screen times both and keeps the
        // faster, running it last so its partials are this call's.
        auto two_launches = [&]() -> int {
            cp.n64 = nullptr;
            for (int seg = 0; seg <= precision; ++seg) {
                set_segment(b, c, precision, seg, cp, b->d_zero.as<int>());
                SF_TRY(launch_grad_class(c, seg, cp, b->d_gout.as<double>(), s));
            }
            return kOk;
        };
        auto one_launch = [&]() -> int {
            set_segment(b, c, 0, 0, cp, b->d_zero.as<int>());
            cp.tau_prim = b->tau_prim64;
            cp.n64 = Q.d_k032.as<int>();
            cp.tau_prim_lo = b->tau_prim;
            return launch_grad_class(c, 0, cp, b->d_gout.as<double>(), s);
        };
        if (precision == 1 && Q.nq32 > 0) {
            if (b->grad_arith[c] < 0) {
                float ms[2] = {0.f, 0.f};
                for (int v = 0; v < 2; ++v) {
                    hipEvent_t e0, e1;
                    SF_HIP(hipEventCreate(&e0));
                    SF_HIP(hipEventCreate(&e1));
                    SF_HIP(hipEventRecord(e0, s));
                    const int rc = v == 0 ? two_launches() : one_launch();
                    SF_HIP(hipEventRecord(e1, s));
                    SF_HIP(hipEventSynchronize(e1));
                    SF_HIP(hipEventElapsedTime(&ms[v], e0, e1));
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    SF_TRY(rc);
                }
                b->grad_arith[c] = ms[1] < ms[0] ? 1 : 0;
                if (b->grad_arith[c] == 0) SF_TRY(two_launches());
            } else if (b->grad_arith[c] == 1) {
                SF_TRY(one_launch());
            } else {
                SF_TRY(two_launches());
            }
        } else {
            SF_TRY(two_launches());
        }
        SF_TRY(launch_grad_reduce_bra(b->d_gout.as<double>(), class_nslice(c), Q.nq_all,
                                      Q.d_base_all.as<long long>(), Q.nbra,
                                      b->d_gpart.as<double>() + b->gpart_bra[c],
                                      b->d_gkc.as<double>() + 6 * b->gkc_q,
                                      b->d_gkc.as<double>(), s));
        SF_TRY(launch_grad_reduce_ket(b->d_gkc.as<double>(), Q.d_base_all.as<long long>(), Q.nbra,
                                      (int)b->pc[Q.pck].s_sa.size(), Q.same,
                                      b->d_gpart.as<double>() + b->gpart_ket[c], s));
    }
    SF_TRY(launch_grad_atoms(b->d_gpart.as<double>(), b->d_gcsr_off.as<long long>(),
                             b->d_gcsr_idx.as<long long>(), natom, b->d_g2.as<double>(), s));
    long long row = 0;
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& Pc = b->pc[p];
        SF_TRY(launch_grad_one_electron(p, Pc.d_sa.as<int>(), Pc.d_sb.as<int>(), Pc.d_poff.as<int>(),
                                        Pc.d_np.as<int>(), (int)Pc.sa.size(),
                                        b->d_prim.as<PrimPair>(), b->d_xyz.as<double>(),
                                        b->map(), b->d_atom.as<int>(), natom,
                                        b->d_Z.as<double>(), b->d_nuc.as<double>(),
                                        b->d_Ps.as<double>(), b->d_Ws.as<double>(),
                                        b->d_grows.as<double>() + row * 3 * natom,
                                        b->d_gonee.as<double>() + row * 6 * natom, s));
        row += (long long)Pc.sa.size();
    }
    SF_TRY(launch_grad_rows(b->d_grows.as<double>(), row, 3 * natom, b->d_g2.as<double>(),
                            b->d_gtot.as<double>(), s));
    SF_HIP(hipMemcpyAsync(grad, b->d_gtot.p, 3 * natom * sizeof(double), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    return kOk;
}
