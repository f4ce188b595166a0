// sf_api.hip -- the C-ABI of libsfeng.so: basis handles, Schwarz screening,
// quartet lists, the deterministic digestion plan and the J/K driver.
//
// The reference's call sites replaced here: rys.eri_rys via f2py
// (src/hf/hf.py:26-30, :100-104), build_eri_parallel (hf.py:33-61) and the
// J/K einsums (hf.py:300).
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
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
    // every pair of the class, canonical orientation (L(sa) >= L(sb))
    std::vector<int> sa, sb, poff, np;
    std::vector<double> Q;
    DevBuf d_sa, d_sb, d_poff, d_np, d_Q;
    // significant pairs, sorted by Q descending
    std::vector<int> s_sa, s_sb, s_poff, s_np;
    std::vector<double> s_Q;
    DevBuf ds_sa, ds_sb, ds_poff, ds_np, ds_Q;
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
    std::vector<int> n_all, n64;
    std::vector<long long> base_all, base64, base32;
    DevBuf d_base_all, d_base64, d_base32, d_k032, d_pos, d_raw;
    int ntile = 0;
};

struct sf_basis {
    int nshell = 0, nbf = 0, natom = 0;
    std::vector<int> L, nprim, pstart, foff, atom;
    std::vector<double> xyz, expn, coef, fnorm;
    std::vector<PrimPair> prim;
    DevBuf d_prim, d_xyz, d_foff, d_fnorm, d_zero;
    // nuclei (point charges) for the one-electron potential
    int nnuc = 0;
    DevBuf d_Z, d_nuc, d_atom;
    // nuclear gradient scratch (built on first use after each sf_screen)
    bool grad_ready = false;
    std::vector<long long> gpart_bra, gpart_ket;  // per class offsets into d_gpart
    DevBuf d_gpart, d_gout, d_gcsr_off, d_gcsr_idx, d_grows, d_Ws, d_g2, d_gtot;
    PairClassData pc[kNumPairClass];
    std::vector<int> pair_slot;  // for pair (a>=b) index: (pc << 24) | position in pc arrays
    bool have_schwarz = false, screened = false;
    double tau = 0.0, tau_mixed = 0.0;
    QuartetClassData qc[kNumQuartetClass];
    long long funcq = 0, primq = 0;
    // digestion plan: destination-sorted tile slots
    DevBuf d_segb, d_sege, d_key_base, d_tiles;
    DevBuf d_dest[9];
    int ndest[9] = {0};
    unsigned long long ntile_total = 0;
    DevBuf d_Ps, d_X, d_Y;
};

extern "C" const char* sf_last_error(void) { return g_err; }
extern "C" int sf_version(void) { return 100; }

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

extern "C" int sf_eri_components(int nf, const int* npgf, const int* poff, const double* coefs,
                                 const double* expns, const double* norms, const int* shell,
                                 const double* center, long long nq, const int* quartets,
                                 double* out)
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
        if (L > 9) {
            set_error(kEUnsupported, "quartet %lld: total angular momentum %d > 9", q, L);
            return kEUnsupported;
        }
    }
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
extern "C" int sf_basis_create(int nshell, const int* L, const int* nprim, const double* center,
                               const double* expns, const double* coefs, int natom,
                               const int* shell_atom, sf_basis** out)
{
    if (!out || nshell <= 0 || !L || !nprim || !center || !expns || !coefs || natom < 0) {
        set_error(kEInval, "sf_basis_create: bad arguments");
        return kEInval;
    }
    *out = nullptr;
    sf_basis* b = new (std::nothrow) sf_basis();
    if (!b) {
        set_error(kENoMem, "sf_basis_create: host allocation");
        return kENoMem;
    }
    b->nshell = nshell;
    b->natom = natom;
    int off = 0, pst = 0;
    for (int s = 0; s < nshell; ++s) {
        if (L[s] < 0 || L[s] > kMaxL) {
            set_error(kEUnsupported, "shell %d: L = %d (s, p, d supported)", s, L[s]);
            delete b;
            return kEUnsupported;
        }
        if (nprim[s] <= 0 || nprim[s] > 64) {
            set_error(kEInval, "shell %d: nprim = %d", s, nprim[s]);
            delete b;
            return kEInval;
        }
        b->L.push_back(L[s]);
        b->nprim.push_back(nprim[s]);
        b->pstart.push_back(pst);
        b->foff.push_back(off);
        b->atom.push_back(shell_atom ? shell_atom[s] : 0);
        for (int k = 0; k < 3; ++k) b->xyz.push_back(center[3 * s + k]);
        for (int i = 0; i < nprim[s]; ++i) {
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
    b->nbf = off;
    // all shell pairs a >= b, canonical orientation, primitive pair data.
    // Primitive-pair screening: with Kp = K/p, a primitive quartet is bounded
    // by 2 pi^{5/2} |Kp_ab| |Kp_cd| / sqrt(p+q) times its polynomial factors;
    // pairs with 2 pi^{5/2} |Kp_ab| M / sqrt(p_ab) < 1e-20 (M = largest |Kp|
    // of the basis) are dropped, which changes no integral by more than ~1e-15.
    std::vector<double> kp_all;
    double kp_max = 0.0;
    for (int a = 0; a < nshell; ++a)
        for (int c = 0; c <= a; ++c) {
            const double* A = &b->xyz[3 * a];
            const double* B = &b->xyz[3 * c];
            const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                               (A[2] - B[2]) * (A[2] - B[2]);
            for (int i = 0; i < nprim[a]; ++i)
                for (int j = 0; j < nprim[c]; ++j) {
                    const double ea = b->expn[b->pstart[a] + i], eb = b->expn[b->pstart[c] + j];
                    const double p = ea + eb;
                    const double K = b->coef[b->pstart[a] + i] * b->coef[b->pstart[c] + j] *
                                     std::exp(-ea * eb / p * AB2);
                    kp_max = std::max(kp_max, std::fabs(K / p));
                }
        }
    const double kPrimTol = 1e-20;
    b->pair_slot.resize((size_t)nshell * (nshell + 1) / 2);
    for (int a = 0; a < nshell; ++a)
        for (int c = 0; c <= a; ++c) {
            int sa = a, sb = c;
            if (L[sb] > L[sa]) std::swap(sa, sb);
            const int pcl = pair_class(L[sa], L[sb]);
            PairClassData& P = b->pc[pcl];
            P.sa.push_back(sa);
            P.sb.push_back(sb);
            P.poff.push_back((int)b->prim.size());
            b->pair_slot[(size_t)a * (a + 1) / 2 + c] = (pcl << 24) | (int)(P.sa.size() - 1);
            const double* A = &b->xyz[3 * sa];
            const double* B = &b->xyz[3 * sb];
            const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                               (A[2] - B[2]) * (A[2] - B[2]);
            int kept = 0;
            for (int i = 0; i < nprim[sa]; ++i)
                for (int j = 0; j < nprim[sb]; ++j) {
                    const double ea = b->expn[b->pstart[sa] + i], eb = b->expn[b->pstart[sb] + j];
                    const double p = ea + eb;
                    PrimPair pp;
                    pp.p = p;
                    pp.Px = (ea * A[0] + eb * B[0]) / p;
                    pp.Py = (ea * A[1] + eb * B[1]) / p;
                    pp.Pz = (ea * A[2] + eb * B[2]) / p;
                    pp.K = b->coef[b->pstart[sa] + i] * b->coef[b->pstart[sb] + j] *
                           std::exp(-ea * eb / p * AB2);
                    pp.ip = 1.0 / p;
                    pp.Kp = pp.K / p;
                    pp.eb = eb;
                    if (34.986836655249725 * std::fabs(pp.Kp) * kp_max / std::sqrt(p) < kPrimTol)
                        continue;
                    b->prim.push_back(pp);
                    ++kept;
                }
            P.np.push_back(kept);
        }
    if (b->pair_slot.size() > (1u << 24)) {
        set_error(kEUnsupported, "too many shell pairs");
        delete b;
        return kEUnsupported;
    }
    int rc = b->d_prim.upload(b->prim);
    if (!rc) rc = b->d_atom.upload(b->atom);
    if (!rc) rc = b->d_xyz.upload(b->xyz);
    if (!rc) rc = b->d_foff.upload(b->foff);
    if (!rc) rc = b->d_fnorm.upload(b->fnorm);
    for (int p = 0; p < kNumPairClass && !rc; ++p) {
        PairClassData& P = b->pc[p];
        rc = P.d_sa.upload(P.sa);
        if (!rc) rc = P.d_sb.upload(P.sb);
        if (!rc) rc = P.d_poff.upload(P.poff);
        if (!rc) rc = P.d_np.upload(P.np);
        if (!rc) rc = P.d_Q.ensure(std::max<size_t>(1, P.sa.size()) * sizeof(double));
    }
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
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

static int compute_schwarz(sf_basis* b)
{
    if (b->have_schwarz) return kOk;
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        const int n = (int)P.sa.size();
        if (!n) continue;
        SF_TRY(launch_schwarz(p, P.d_sa.as<int>(), P.d_sb.as<int>(), P.d_poff.as<int>(),
                              P.d_np.as<int>(), n, b->d_prim.as<PrimPair>(), b->d_xyz.as<double>(),
                              P.d_Q.as<double>(), 0));
    }
    SF_HIP(hipDeviceSynchronize());
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        P.Q.resize(P.sa.size());
        if (!P.Q.empty())
            SF_HIP(hipMemcpy(P.Q.data(), P.d_Q.p, P.Q.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    b->have_schwarz = true;
    return kOk;
}

extern "C" int sf_schwarz(sf_basis* b, double* Q)
{
    if (!b || !Q) {
        set_error(kEInval, "sf_schwarz: null argument");
        return kEInval;
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

extern "C" int sf_screen(sf_basis* b, double tau, double tau_mixed)
{
    if (!b || !(tau >= 0.0) || !(tau_mixed >= 0.0)) {
        set_error(kEInval, "sf_screen: bad arguments");
        return kEInval;
    }
    SF_TRY(compute_schwarz(b));
    b->screened = false;
    b->grad_ready = false;
    b->tau = tau;
    b->tau_mixed = tau_mixed;
    double qmax = 0.0;
    for (auto& P : b->pc)
        for (double q : P.Q) qmax = std::max(qmax, q);
    // significant pairs, sorted by Q descending (ties: original order)
    for (auto& P : b->pc) {
        std::vector<int> idx;
        for (int i = 0; i < (int)P.Q.size(); ++i)
            if (P.Q[i] * qmax >= tau) idx.push_back(i);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return P.Q[x] > P.Q[y]; });
        P.s_sa.clear(); P.s_sb.clear(); P.s_poff.clear(); P.s_np.clear(); P.s_Q.clear();
        for (int i : idx) {
            P.s_sa.push_back(P.sa[i]);
            P.s_sb.push_back(P.sb[i]);
            P.s_poff.push_back(P.poff[i]);
            P.s_np.push_back(P.np[i]);
            P.s_Q.push_back(P.Q[i]);
        }
        SF_TRY(P.ds_sa.upload(P.s_sa));
        SF_TRY(P.ds_sb.upload(P.s_sb));
        SF_TRY(P.ds_poff.upload(P.s_poff));
        SF_TRY(P.ds_np.upload(P.s_np));
        SF_TRY(P.ds_Q.upload(P.s_Q));
    }
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
            if (Q.nbra > 0 && !K.s_sa.empty()) {
                SF_TRY(launch_count_kets(B.ds_Q.as<double>(), Q.nbra, K.ds_Q.as<double>(),
                                         (int)K.s_sa.size(), Q.same, tau, tau_mixed,
                                         d_nall.as<int>(), d_n64.as<int>(), 0));
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
            SF_TRY(Q.d_k032.upload(Q.n64));
            SF_TRY(Q.d_pos.ensure((size_t)std::max(1LL, Q.nq_all) * 6 * sizeof(unsigned)));
            if (class_split(c))
                SF_TRY(Q.d_raw.ensure((size_t)std::max(1LL, Q.nq_all) * class_nint(c) * sizeof(double)));
            else
                Q.d_raw.release();
            total_entries += 6 * Q.nq_all;
            // function-level unique quartets and primitive quartets
            const int la = pair_la(pcb), lb = pair_lb(pcb), lc = pair_la(pck), ld = pair_lb(pck);
            const long long nA = cart_n(la), nB = cart_n(lb), nC = cart_n(lc), nD = cart_n(ld);
            for (int i = 0; i < Q.nbra; ++i) {
                const bool eqab = B.s_sa[i] == B.s_sb[i];
                const long long pab = eqab ? nA * (nA + 1) / 2 : nA * nB;
                for (int j = 0; j < Q.n_all[i]; ++j) {
                    const bool eqcd = K.s_sa[j] == K.s_sb[j];
                    const long long pcd = eqcd ? nC * (nC + 1) / 2 : nC * nD;
                    b->funcq += (Q.same && i == j) ? pab * (pab + 1) / 2 : pab * pcd;
                    b->primq += (long long)B.s_np[i] * K.s_np[j];
                }
            }
        }
    // digestion plan: (destination, entry) pairs sorted by destination
    const int ns = b->nshell;
    const unsigned long long nkeys = 2ULL * ns * ns;
    if (nkeys >= (1ULL << 32)) {
        set_error(kEUnsupported, "too many shells for the digestion plan");
        return kEUnsupported;
    }
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
                                   Q.d_base_all.as<long long>(), Q.nq_all, ns,
                                   d_keys.as<unsigned>() + ebase,
                                   d_vals.as<unsigned long long>() + ebase, 0));
        ebase += 6 * Q.nq_all;
    }
    DevBuf d_tab;
    SF_TRY(d_tab.upload(tab));
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
    // slot layout: destinations in key order, each a run of NE-double tiles
    std::vector<unsigned long long> key_base(nkeys);
    std::vector<int> key_ne(nkeys);
    std::vector<int> dest[9];
    unsigned long long acc = 0;
    const unsigned long long nsq = (unsigned long long)ns * ns;
    for (unsigned long long k = 0; k < nkeys; ++k) {
        const unsigned long long rem = k % nsq;
        const int s1 = (int)(rem / ns), s2 = (int)(rem % ns);
        key_ne[k] = cart_n(b->L[s1]) * cart_n(b->L[s2]);
        key_base[k] = acc;
        if (sege[k] <= segb[k]) continue;
        acc += (unsigned long long)(sege[k] - segb[k]) * key_ne[k];
        dest[3 * b->L[s1] + b->L[s2]].push_back((int)k);
    }
    if (acc >= (1ULL << 32)) {
        set_error(kEUnsupported, "digestion plan of %llu doubles exceeds 32-bit slots", acc);
        return kEUnsupported;
    }
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
    for (int g = 0; g < 9; ++g) {
        b->ndest[g] = (int)dest[g].size();
        SF_TRY(b->d_dest[g].upload(dest[g]));
    }
    const size_t nn = (size_t)b->nbf * b->nbf;
    SF_TRY(b->d_Ps.ensure(nn * sizeof(double)));
    SF_TRY(b->d_X.ensure(nn * sizeof(double)));
    SF_TRY(b->d_Y.ensure(nn * sizeof(double)));
    SF_HIP(hipDeviceSynchronize());
    b->screened = true;
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

extern "C" int sf_pair_list(const sf_basis* b, int pc, int* n, int* sa, int* sb, double* Q,
                            int* nprim)
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

static ClassParams class_params(sf_basis* b, int c, const double* Ps)
{
    QuartetClassData& Q = b->qc[c];
    ClassParams cp;
    cp.bra = b->pc[Q.pcb].sorted_dev();
    cp.ket = b->pc[Q.pck].sorted_dev();
    cp.prim = b->d_prim.as<PrimPair>();
    cp.shell_xyz = b->d_xyz.as<double>();
    cp.shell_foff = b->d_foff.as<int>();
    cp.seg_base = Q.d_base_all.as<long long>();
    cp.seg_k0 = b->d_zero.as<int>();
    cp.qall_base = Q.d_base_all.as<long long>();
    cp.nq = Q.nq_all;
    cp.qstride = Q.nq_all;
    cp.same = Q.same;
    cp.nbf = b->nbf;
    cp.P = Ps;
    cp.out = nullptr;
    cp.raw = Q.d_raw.as<double>();
    cp.pos = Q.d_pos.as<unsigned>();
    cp.tiles = b->d_tiles.as<double>();
    return cp;
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
        cp.out = d_dense.as<double>();
        SF_TRY(launch_eri_class(c, 0, kModeRaw, cp, 0));
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

// J/K driver; when ms is non-null, HIP events bracket every launch group on
// the stream and ms receives milliseconds per group: [2c] FP64 and [2c+1]
// FP32 ERI+digestion kernels of quartet class c (c = 0..20), [42] density
// scaling, [43] plan reduction + finalisation.
static int run_jk(sf_basis* b, const double* P, double* J, double* K, int precision,
                  hipStream_t s, double* ms)
{
    if (!b || !P || !J || !K || precision < 0 || precision > 1) {
        set_error(kEInval, "sf_jk: bad arguments");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_jk: call sf_screen first");
        return kEState;
    }
    const int nbf = b->nbf;
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
    SF_TRY(launch_scale_density(P, b->d_fnorm.as<double>(), nbf, b->d_Ps.as<double>(), s));
    SF_TRY(mark());
    slot.push_back(42);
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        if (Q.nq_all == 0) continue;
        ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        if (precision == 0) {
            SF_TRY(launch_eri_class(c, 0, kModeJK, cp, s));
            SF_TRY(mark());
            slot.push_back(2 * c);
        } else {
            cp.seg_base = Q.d_base64.as<long long>();
            cp.seg_k0 = b->d_zero.as<int>();
            cp.nq = Q.nq64;
            SF_TRY(launch_eri_class(c, 0, kModeJK, cp, s));
            SF_TRY(mark());
            slot.push_back(2 * c);
            cp.seg_base = Q.d_base32.as<long long>();
            cp.seg_k0 = Q.d_k032.as<int>();
            cp.nq = Q.nq32;
            SF_TRY(launch_eri_class(c, 1, kModeJK, cp, s));
            SF_TRY(mark());
            slot.push_back(2 * c + 1);
        }
    }
    const long long nn = (long long)nbf * nbf;
    SF_TRY(launch_fill(b->d_X.as<double>(), nn, 0.0, s));
    SF_TRY(launch_fill(b->d_Y.as<double>(), nn, 0.0, s));
    const int nf[3] = {1, 3, 6};
    for (int g = 0; g < 9; ++g)
        SF_TRY(launch_reduce(nf[g / 3], nf[g % 3], b->d_dest[g].as<int>(), b->ndest[g],
                             b->d_segb.as<long long>(), b->d_sege.as<long long>(),
                             b->d_key_base.as<unsigned long long>(), b->d_tiles.as<double>(),
                             b->d_foff.as<int>(), b->nshell, nbf, b->d_X.as<double>(),
                             b->d_Y.as<double>(), s));
    SF_TRY(launch_finalize(b->d_X.as<double>(), b->d_Y.as<double>(), b->d_fnorm.as<double>(), nbf,
                           J, K, s));
    SF_TRY(mark());
    slot.push_back(43);
    if (ms) {
        for (int k = 0; k < 44; ++k) ms[k] = 0.0;
        SF_HIP(hipEventSynchronize(ev.back()));
        for (size_t k = 1; k < ev.size(); ++k) {
            float t = 0.f;
            SF_HIP(hipEventElapsedTime(&t, ev[k - 1], ev[k]));
            ms[slot[k - 1]] += t;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
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

extern "C" int sf_basis_set_nuclei(sf_basis* b, int natom, const double* Z, const double* xyz)
{
    if (!b || natom < 0 || (natom > 0 && (!Z || !xyz))) {
        set_error(kEInval, "sf_basis_set_nuclei: bad arguments");
        return kEInval;
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
    for (int p = 0; p < kNumPairClass; ++p) {
        PairClassData& P = b->pc[p];
        SF_TRY(launch_one_electron(p, P.d_sa.as<int>(), P.d_sb.as<int>(), P.d_poff.as<int>(),
                                   P.d_np.as<int>(), (int)P.sa.size(), b->d_prim.as<PrimPair>(),
                                   b->d_xyz.as<double>(), b->d_foff.as<int>(),
                                   b->d_fnorm.as<double>(), b->nnuc, b->d_Z.as<double>(),
                                   b->d_nuc.as<double>(), b->nbf, S, T, V, (hipStream_t)stream));
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
    const size_t nn = (size_t)b->nbf * b->nbf;
    SF_TRY(b->d_Ws.ensure(nn * sizeof(double)));
    SF_TRY(b->d_g2.ensure(3 * std::max(1, natom) * sizeof(double)));
    SF_TRY(b->d_gtot.ensure(3 * std::max(1, natom) * sizeof(double)));
    b->grad_ready = true;
    return kOk;
}

extern "C" int sf_grad(sf_basis* b, const double* P, const double* W, double* grad, void* stream)
{
    if (!b || !P || !W || !grad) {
        set_error(kEInval, "sf_grad: null argument");
        return kEInval;
    }
    if (!b->screened) {
        set_error(kEState, "sf_grad: call sf_screen first");
        return kEState;
    }
    if (b->nnuc != b->natom || b->natom <= 0) {
        set_error(kEState, "sf_grad: nuclei (sf_basis_set_nuclei) must match the %d atoms", b->natom);
        return kEState;
    }
    hipStream_t s = (hipStream_t)stream;
    SF_TRY(prepare_grad(b));
    const int nbf = b->nbf, natom = b->natom;
    SF_TRY(launch_scale_density(P, b->d_fnorm.as<double>(), nbf, b->d_Ps.as<double>(), s));
    SF_TRY(launch_scale_density(W, b->d_fnorm.as<double>(), nbf, b->d_Ws.as<double>(), s));
    for (int c = 0; c < kNumQuartetClass; ++c) {
        QuartetClassData& Q = b->qc[c];
        if (Q.nq_all == 0) continue;
        ClassParams cp = class_params(b, c, b->d_Ps.as<double>());
        SF_TRY(launch_grad_class(c, cp, b->d_gout.as<double>(), s));
        SF_TRY(launch_grad_reduce_bra(b->d_gout.as<double>(), class_nslice(c), Q.nq_all,
                                      Q.d_base_all.as<long long>(), Q.nbra,
                                      b->d_gpart.as<double>() + b->gpart_bra[c], s));
        SF_TRY(launch_grad_reduce_ket(b->d_gout.as<double>(), class_nslice(c), Q.nq_all,
                                      Q.d_base_all.as<long long>(), Q.nbra,
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
                                        b->d_foff.as<int>(), b->d_atom.as<int>(), natom,
                                        b->d_Z.as<double>(), b->d_nuc.as<double>(), nbf,
                                        b->d_Ps.as<double>(), b->d_Ws.as<double>(),
                                        b->d_grows.as<double>() + row * 3 * natom, s));
        row += (long long)Pc.sa.size();
    }
    SF_TRY(launch_grad_rows(b->d_grows.as<double>(), row, 3 * natom, b->d_g2.as<double>(),
                            b->d_gtot.as<double>(), s));
    SF_HIP(hipMemcpyAsync(grad, b->d_gtot.p, 3 * natom * sizeof(double), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    return kOk;
}
