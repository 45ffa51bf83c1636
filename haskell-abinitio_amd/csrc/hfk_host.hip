// Host side of the C ABI (include/hfk.h): context construction (shells,
// primitive pairs, Boys tables, Schwarz bounds) and the integral-direct
// Fock build.  All arithmetic runs in HIP kernels; this file only plans
// and launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "hfk.h"
#include "hfk_internal.hpp"
#include "hfk_host.hpp"

static thread_local std::string g_last_error;

int hfk_set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" int hfk_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return HFK_EINVAL;
    snprintf(buf, len, "%s", g_last_error.c_str());
    return HFK_OK;
}

extern "C" void hfk_default_options(hfk_options* opt)
{
    opt->screen_mode = HFK_SCREEN_REFERENCE;
    opt->boys_mode = HFK_BOYS_REFERENCE;
    opt->screen_tau = 1e-8;     // IntegralsEvaluation.hs:309
    opt->schwarz_cut = 0.0;
    opt->fp32_threshold = 0.0;
}

namespace {

int ncart(int l) { return (l + 1) * (l + 2) / 2; }

int cart_index(int l, int x, int y, int z)
{
    int k = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (i == x && j == y && (l - i - j) == z) return k;
            ++k;
        }
    return -1;
}

int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }   // la >= lb

// Gauss-Legendre 256: positive nodes ascending, by Newton on P_256 in long
// double (the values of NumericalTools/PointsWeights.hs:25-31).
void gl256(double* x, double* w)
{
    const int N = 256;
    for (int i = 0; i < N / 2; i++) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (N + 0.5L));
        long double pp = 0.0L;
        for (int it = 0; it < 100; it++) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= N; j++) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = N * (z * p1 - p2) / (z * z - 1.0L);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) < 1e-19L) break;
        }
        long double p1 = 1.0L, p2 = 0.0L;
        for (int j = 1; j <= N; j++) {
            long double p3 = p2;
            p2 = p1;
            p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
        }
        pp = N * (z * p1 - p2) / (z * z - 1.0L);
        x[N / 2 - 1 - i] = (double)z;
        w[N / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
    }
}

// facOdd (Boys.hs:89-94) with Int (64-bit wrapping) factorials (fac, Boys.hs:97-100).
double hs_facodd(int i)
{
    if (i < 2) return 1.0;
    int k = (1 + i) / 2;
    uint64_t f2k = 1, fk = 1;
    for (int t = 1; t <= 2 * k; ++t) f2k *= (uint64_t)t;
    for (int t = 1; t <= k; ++t) fk *= (uint64_t)t;
    double p2 = 1.0;
    for (int t = 0; t < k; ++t) p2 *= 2.0;
    return (double)(int64_t)f2k / (p2 * (double)(int64_t)fk);
}

}  // namespace

// ----------------------------------------------------------------------
// context construction
// ----------------------------------------------------------------------

static int group_shells(hfk_ctx* c, const hfk_basis* b)
{
    const int n = b->nbf;
    int i = 0;
    while (i < n) {
        const int* l = b->lxyz + 3 * i;
        const int L = l[0] + l[1] + l[2];
        if (L > HFK_LMAX || l[0] < 0 || l[1] < 0 || l[2] < 0)
            return hfk_set_error(HFK_EUNSUP, "basis function %d has angular momentum %d (max %d)", i, L, HFK_LMAX);
        HostShell s;
        s.atom = b->atom_of_bf ? b->atom_of_bf[i] : -1;
        for (int x = 0; x < 3; ++x) s.R[x] = b->center[3 * i + x];
        s.L = L;
        const int p0 = b->prim_off[i], np = b->prim_off[i + 1] - p0;
        if (np <= 0) return hfk_set_error(HFK_EINVAL, "basis function %d has no primitives", i);
        s.prim0 = (int)c->sh_expo.size();
        s.nprim = np;
        for (int k = 0; k < np; ++k) {
            c->sh_expo.push_back(b->expo[p0 + k]);
            c->sh_coef.push_back(b->coef[p0 + k]);
        }
        for (int k = 0; k < HFK_MAXCOMP; ++k) { s.bf[k] = -1; s.scale[k] = 0.0; }
        int ci = cart_index(L, l[0], l[1], l[2]);
        s.bf[ci] = i;
        s.scale[ci] = 1.0;
        int j = i + 1;
        for (; j < n; ++j) {
            const int* lj = b->lxyz + 3 * j;
            if (lj[0] + lj[1] + lj[2] != L) break;
            bool same = b->center[3 * j] == s.R[0] && b->center[3 * j + 1] == s.R[1] && b->center[3 * j + 2] == s.R[2];
            if (!same) break;
            const int q0 = b->prim_off[j], nq = b->prim_off[j + 1] - q0;
            if (nq != np) break;
            bool ok = true;
            for (int k = 0; k < np && ok; ++k) ok = b->expo[q0 + k] == b->expo[p0 + k];
            if (!ok) break;
            const int cj = cart_index(L, lj[0], lj[1], lj[2]);
            if (s.bf[cj] >= 0) break;
            // coefficients must be proportional (component normalisation)
            const double r = b->coef[q0] / b->coef[p0];
            for (int k = 0; k < np && ok; ++k) {
                const double rk = b->coef[q0 + k] / b->coef[p0 + k];
                ok = fabs(rk - r) <= 1e-13 * fabs(r);
            }
            if (!ok) break;
            s.bf[cj] = j;
            s.scale[cj] = r;
        }
        c->shells.push_back(s);
        i = j;
    }
    return HFK_OK;
}

int hfk_ctx_build_pairs(hfk_ctx* c);
int hfk_ctx_schwarz(hfk_ctx* c);

extern "C" int hfk_create(const hfk_basis* b, const hfk_options* opt, hfk_ctx** out)
{
    if (!b || !out || b->nbf <= 0 || !b->center || !b->lxyz || !b->prim_off || !b->coef || !b->expo)
        return hfk_set_error(HFK_EINVAL, "hfk_create: null or empty basis");
    *out = nullptr;
    hfk_ctx* c = new hfk_ctx();
    if (opt) c->opt = *opt;
    else hfk_default_options(&c->opt);
    c->n = b->nbf;
    c->natom = b->natom;
    c->center.assign(b->center, b->center + 3 * (size_t)b->nbf);
    c->lxyz.assign(b->lxyz, b->lxyz + 3 * (size_t)b->nbf);
    c->prim_off.assign(b->prim_off, b->prim_off + b->nbf + 1);
    const int nprim = b->prim_off[b->nbf];
    c->coef.assign(b->coef, b->coef + nprim);
    c->expo.assign(b->expo, b->expo + nprim);
    if (b->natom > 0 && b->Z && b->R) {
        c->Z.assign(b->Z, b->Z + b->natom);
        c->R.assign(b->R, b->R + 3 * (size_t)b->natom);
    }
    if (b->atom_of_bf) c->atom_of_bf.assign(b->atom_of_bf, b->atom_of_bf + b->nbf);
    int rc = group_shells(c, b);
    if (rc == HFK_OK) rc = hfk_ctx_init_device(c);
    if (rc == HFK_OK) rc = hfk_ctx_build_pairs(c);
    if (rc == HFK_OK) rc = hfk_ctx_schwarz(c);
    if (rc == HFK_OK) rc = hfk_ctx_plan(c);
    if (rc != HFK_OK) {
        hfk_ctx_free_device(c);
        delete c;
        return rc;
    }
    *out = c;
    return HFK_OK;
}

extern "C" int hfk_destroy(hfk_ctx* c)
{
    if (!c) return HFK_OK;
    hfk_ctx_free_device(c);
    delete c;
    return HFK_OK;
}

extern "C" int hfk_set_options(hfk_ctx* c, const hfk_options* opt)
{
    if (!c || !opt) return hfk_set_error(HFK_EINVAL, "hfk_set_options: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    const bool rebuild = opt->boys_mode != c->opt.boys_mode || opt->screen_mode != c->opt.screen_mode ||
                         opt->screen_tau != c->opt.screen_tau;
    c->opt = *opt;
    c->scf_cached = false;   // its Fock matrix was built under the old options
    int rc = HFK_OK;
    if (rebuild) {
        rc = hfk_ctx_classify_pairs(c);
        if (rc == HFK_OK) rc = hfk_ctx_schwarz(c);
    }
    if (rc == HFK_OK) rc = hfk_ctx_plan(c);
    return rc;
}

// Idle streams of destroyed contexts, per device.  Creating a stream costs
// about 0.7 ms on this runtime and a context holds four; a caller that makes
// many short-lived contexts (the fragment guess makes one per monomer) takes
// them from here instead.  Streams are returned synchronised; at most
// kStreamPoolMax are kept, the rest destroyed.
namespace {
constexpr size_t kStreamPoolMax = 64;
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_stream_pool;
}  // namespace

static hipError_t stream_acquire(int dev, hipStream_t* s)
{
    {
        std::lock_guard<std::mutex> g(g_stream_mu);
        for (size_t i = g_stream_pool.size(); i-- > 0;)
            if (g_stream_pool[i].first == dev) {
                *s = g_stream_pool[i].second;
                g_stream_pool.erase(g_stream_pool.begin() + (long)i);
                return hipSuccess;
            }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

static void stream_release(int dev, hipStream_t s)
{
    if (!s) return;
    if (hipStreamSynchronize(s) == hipSuccess) {
        std::lock_guard<std::mutex> g(g_stream_mu);
        if (g_stream_pool.size() < kStreamPoolMax) {
            g_stream_pool.emplace_back(dev, s);
            return;
        }
    }
    hipStreamDestroy(s);
}

int hfk_ctx_init_device(hfk_ctx* c)
{
    HFK_HIP(hipGetDevice(&c->device));
    HFK_TRY(hfk_noalloc_setup(c->device));
    HFK_HIP(stream_acquire(c->device, &c->stream));
    for (int k = 0; k < hfk_ctx::NAUX; ++k) {
        HFK_HIP(stream_acquire(c->device, &c->aux[k]));
        HFK_HIP(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
    }
    HFK_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (const char* v = getenv("HFK_CONCURRENT")) c->concurrent = atoi(v) != 0;
    HFK_HIP(hipHostMalloc((void**)&c->h_pin, sizeof(double) * HFK_PIN_DOUBLES + sizeof(int), hipHostMallocDefault));
    // Boys tables
    std::vector<double> glx(128), glw(128);
    gl256(glx.data(), glw.data());
    HFK_TRY(dev_upload(c, &c->d_glx, glx));
    HFK_TRY(dev_upload(c, &c->d_glw, glw));
    HFK_TRY(dev_alloc(&c->d_boys_ref64, (size_t)BOYS_REF_STRIDE * BOYS_NX));
    HFK_TRY(dev_alloc(&c->d_boys_ref32, (size_t)BOYS_REF_STRIDE * BOYS_NX));
    HFK_TRY(dev_alloc(&c->d_boys_acc64, (size_t)BOYS_ACC_STRIDE * BOYS_ACC_NX));
    HFK_TRY(dev_alloc(&c->d_boys_acc32, (size_t)BOYS_ACC_STRIDE * BOYS_ACC_NX));
    hfk_launch_fill(c->d_boys_ref64, 0.0, (long long)BOYS_REF_STRIDE * BOYS_NX, c->stream);
    hfk_launch_fill(c->d_boys_acc64, 0.0, (long long)BOYS_ACC_STRIDE * BOYS_ACC_NX, c->stream);
    hfk_launch_boys_tables(c->d_glx, c->d_glw, c->d_boys_ref64, c->d_boys_ref32, c->d_boys_acc64, c->d_boys_acc32,
                           c->stream);
    HFK_HIP(hipGetLastError());
    c->bt.ref64 = c->d_boys_ref64;
    c->bt.ref32 = c->d_boys_ref32;
    c->bt.acc64 = c->d_boys_acc64;
    c->bt.acc32 = c->d_boys_acc32;
    for (int m = 0; m < 24; ++m) {
        // asymptBoysF m x = (facOdd (floor (2m-1)) / 2^(m+1)) * sqrt (pi / x^(2m+1))  (Boys.hs:85-87)
        c->bt.asym[m] = hs_facodd((int)floor(2.0 * m - 1.0)) / pow(2.0, m + 1.0);
    }
    // raw basis on device (one-electron integrals)
    HFK_TRY(dev_upload(c, &c->d_center, c->center));
    HFK_TRY(dev_upload(c, &c->d_lxyz, c->lxyz));
    HFK_TRY(dev_upload(c, &c->d_prim_off, c->prim_off));
    {
        // upper-triangle function pairs (flat index) grouped by l_i + l_j
        std::vector<int> groups[5];
        const int n = c->n;
        long long idx = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j, ++idx) {
                const int ls = c->lxyz[3 * i] + c->lxyz[3 * i + 1] + c->lxyz[3 * i + 2] + c->lxyz[3 * j] +
                               c->lxyz[3 * j + 1] + c->lxyz[3 * j + 2];
                if (ls > 4) return hfk_set_error(HFK_EUNSUP, "one-electron integrals: l_i + l_j = %d > 4", ls);
                groups[ls].push_back((int)idx);
            }
        std::vector<int> all;
        c->vpair_off.assign(6, 0);
        for (int ls = 0; ls < 5; ++ls) {
            c->vpair_off[ls] = (int)all.size();
            all.insert(all.end(), groups[ls].begin(), groups[ls].end());
        }
        c->vpair_off[5] = (int)all.size();
        HFK_TRY(dev_upload(c, &c->d_vpair, all));
    }
    HFK_TRY(dev_upload(c, &c->d_coef, c->coef));
    HFK_TRY(dev_upload(c, &c->d_expo, c->expo));
    if (!c->Z.empty()) {
        HFK_TRY(dev_upload(c, &c->d_Z, c->Z));
        HFK_TRY(dev_upload(c, &c->d_R, c->R));
    }
    // shells on device
    const int ns = (int)c->shells.size();
    std::vector<double> shc(3 * (size_t)ns), shs((size_t)ns * HFK_MAXCOMP);
    std::vector<int> shL(ns), shp0(ns), shnp(ns), shbf((size_t)ns * HFK_MAXCOMP);
    for (int s = 0; s < ns; ++s) {
        const HostShell& h = c->shells[s];
        for (int x = 0; x < 3; ++x) shc[3 * s + x] = h.R[x];
        shL[s] = h.L;
        shp0[s] = h.prim0;
        shnp[s] = h.nprim;
        for (int k = 0; k < HFK_MAXCOMP; ++k) {
            shbf[(size_t)s * HFK_MAXCOMP + k] = h.bf[k];
            shs[(size_t)s * HFK_MAXCOMP + k] = h.scale[k];
        }
    }
    HFK_TRY(dev_upload(c, &c->d_sh_center, shc));
    HFK_TRY(dev_upload(c, &c->d_sh_L, shL));
    HFK_TRY(dev_upload(c, &c->d_sh_prim0, shp0));
    HFK_TRY(dev_upload(c, &c->d_sh_nprim, shnp));
    HFK_TRY(dev_upload(c, &c->d_sh_bf, shbf));
    std::vector<int> shbfp(shbf);
    for (int& v : shbfp)
        if (v < 0) v = c->n;
    HFK_TRY(dev_upload(c, &c->d_sh_bfp, shbfp));
    HFK_TRY(dev_upload(c, &c->d_sh_scale, shs));
    HFK_TRY(dev_upload(c, &c->d_sh_expo, c->sh_expo));
    HFK_TRY(dev_upload(c, &c->d_sh_coef, c->sh_coef));
    // work matrices
    const size_t nf = (size_t)c->n * (c->n + 1) / 2;
    const size_t np1 = (size_t)(c->n + 1) * (c->n + 1);   // padded: zero row/column n
    HFK_TRY(dev_alloc(&c->d_D, np1));
    HFK_TRY(dev_alloc(&c->d_flatA, nf));
    HFK_TRY(dev_alloc(&c->d_flatB, nf));
    HFK_TRY(dev_alloc(&c->d_flatC, nf));
    HFK_HIP(hipEventCreate(&c->ev0));
    HFK_HIP(hipEventCreate(&c->ev1));
    return HFK_OK;
}

void hfk_ctx_free_device(hfk_ctx* c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) hipFree(p);
    c->allocs.clear();
    hfk_plan_free_arrays(c);
    if (c->h_pin) hipHostFree(c->h_pin);
    c->h_pin = nullptr;
    for (hipEvent_t ev : c->events) hipEventDestroy(ev);
    c->events.clear();
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (int k = 0; k < hfk_ctx::NAUX; ++k) {
        stream_release(c->device, c->aux[k]);
        if (c->ev_join[k]) hipEventDestroy(c->ev_join[k]);
        c->aux[k] = nullptr;
        c->ev_join[k] = nullptr;
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    c->ev_fork = nullptr;
    if (c->blas) hfk_linalg_destroy(c);
    stream_release(c->device, c->stream);
    c->stream = nullptr;
}

// Shell pairs: every (A >= B) by shell index, stored with the shell of the
// larger angular momentum first.  Primitive-pair records and the
// reference's screen values come from k_pair_setup.
int hfk_ctx_build_pairs(hfk_ctx* c)
{
    const int ns = (int)c->shells.size();
    std::vector<int> pA, pB, pp0, npp, q0;
    long long npp_tot = 0, nq_tot = 0;
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b <= a; ++b) {
            int A = a, B = b;
            if (c->shells[B].L > c->shells[A].L) std::swap(A, B);
            pA.push_back(A);
            pB.push_back(B);
            pp0.push_back((int)npp_tot);
            const int k = c->shells[A].nprim * c->shells[B].nprim;
            // tasks carry primitive-pair ranges in 8-bit fields, where 255
            // marks "all of the ket's pairs": at most 254 per pair
            if (k > 254)
                return hfk_set_error(HFK_EUNSUP, "shell pair with %d primitive pairs (at most 254: %d x %d primitives)", k,
                                     c->shells[A].nprim, c->shells[B].nprim);
            npp.push_back(k);
            q0.push_back((int)nq_tot);
            npp_tot += k;
            nq_tot += (long long)k * ncart(c->shells[A].L) * ncart(c->shells[B].L);
        }
    if (npp_tot > INT32_MAX || nq_tot > INT32_MAX)
        return hfk_set_error(HFK_EUNSUP, "too many primitive pairs (%lld)", npp_tot);
    c->npair = (int)pA.size();
    c->h_pair_A = pA;
    c->h_pair_B = pB;
    c->h_pair_npp = npp;
    HFK_TRY(dev_upload(c, &c->d_pair_A, pA));
    HFK_TRY(dev_upload(c, &c->d_pair_B, pB));
    HFK_TRY(dev_upload(c, &c->d_pair_pp0, pp0));
    HFK_TRY(dev_upload(c, &c->d_pair_npp, npp));
    std::vector<int> q0v(q0.begin(), q0.end());
    int* d_q0 = nullptr;
    HFK_TRY(dev_upload(c, &d_q0, q0v));
    HFK_TRY(dev_alloc(&c->d_pp, (size_t)npp_tot * PP_STRIDE));
    HFK_TRY(dev_alloc(&c->d_qcomp, (size_t)std::max(1LL, nq_tot)));
    HFK_TRY(dev_alloc(&c->d_pair_qhi, (size_t)c->npair));
    HFK_TRY(dev_alloc(&c->d_pair_sch, (size_t)c->npair));
    HFK_TRY(dev_alloc(&c->d_pair_diag, (size_t)c->npair));
    ShellDev sd;
    sd.center = c->d_sh_center;
    sd.L = c->d_sh_L;
    sd.prim0 = c->d_sh_prim0;
    sd.nprim = c->d_sh_nprim;
    sd.expo = c->d_sh_expo;
    sd.coef = c->d_sh_coef;
    sd.scale = c->d_sh_scale;
    hfk_launch_pair_setup(sd, c->npair, c->d_pair_A, c->d_pair_B, c->d_pair_pp0, d_q0, c->d_pp, c->d_qcomp,
                          c->d_pair_qhi, c->stream);
    HFK_HIP(hipGetLastError());
    // records of each pair by descending largest screen value
    {
        double *sorted = nullptr, *qsorted = nullptr;
        HFK_TRY(dev_alloc(&sorted, (size_t)npp_tot * PP_STRIDE));
        HFK_TRY(dev_alloc(&qsorted, (size_t)std::max(1LL, nq_tot)));
        hfk_launch_pair_sort(c->npair, c->d_pair_pp0, c->d_pair_npp, d_q0, nq_tot, c->d_pp, c->d_qcomp, sorted, qsorted,
                             c->stream);
        HFK_HIP(hipGetLastError());
        HFK_HIP(hipStreamSynchronize(c->stream));
        dev_free(c, c->d_pp);
        dev_free(c, c->d_qcomp);
        c->d_pp = sorted;
        c->d_qcomp = qsorted;
    }
    c->stats.nshell = ns;
    c->stats.nprim_pair = npp_tot;
    return hfk_ctx_classify_pairs(c);
}

QuartetArgs hfk_ctx_args(hfk_ctx* c)
{
    QuartetArgs a;
    memset(&a, 0, sizeof a);
    a.pp = c->d_pp;
    a.qcomp = c->d_qcomp;
    a.pair_pp0 = c->d_pair_pp0;
    a.pair_npp = c->d_pair_npp;
    a.pair_A = c->d_pair_A;
    a.pair_B = c->d_pair_B;
    a.pair_qhi = c->d_pair_qhi;
    a.pair_sch = c->d_pair_sch;
    a.sh_bf = c->d_sh_bf;
    a.sh_bfp = c->d_sh_bfp;
    a.sh_scale = c->d_sh_scale;
    a.bt = c->bt;
    a.qtau = (c->opt.screen_mode == HFK_SCREEN_REFERENCE) ? c->opt.screen_tau : 0.0;
    a.n = c->n;
    return a;
}

// True Schwarz bounds: sqrt(max_ab |(ab|ab)|) per shell pair, unscreened.
int hfk_ctx_schwarz(hfk_ctx* c)
{
    QuartetArgs a = hfk_ctx_args(c);
    a.qtau = 0.0;
    a.diag = c->d_pair_diag;
    hfk_launch_fill(c->d_pair_diag, 0.0, c->npair, c->stream);
    for (int k = 0; k < HFK_NPC; ++k) {
        if (c->class_count[k] == 0) continue;
        int la = 0, lb = 0;
        hfk_pair_class_l(k, &la, &lb);
        a.bra_list = c->d_class[k];
        a.ket_list = c->d_class[k];
        a.nbra = a.nket = c->class_count[k];
        a.nquart = c->class_count[k];
        int rc = hfk_launch_quartets(la, lb, la, lb, HFK_OP_DIAG, 0, c->opt.boys_mode, 0, a, c->stream);
        if (rc) return hfk_set_error(HFK_EUNSUP, "no kernel for class (%d%d|%d%d)", la, lb, la, lb);
    }
    hfk_launch_sqrt(c->npair, c->d_pair_diag, c->d_pair_sch, c->stream);
    HFK_HIP(hipGetLastError());
    HFK_HIP(hipStreamSynchronize(c->stream));
    return HFK_OK;
}

void hfk_pair_class_l(int k, int* la, int* lb)
{
    int a = 0;
    while ((a + 1) * (a + 2) / 2 <= k) ++a;
    *la = a;
    *lb = k - a * (a + 1) / 2;
}

// Device arrays of one plan entry (freed when the plan is rebuilt).
template <typename T>
static int entry_upload(hfk_ctx* c, PlanEntry& e, T** p, const std::vector<T>& v)
{
    (void)e;
    int rc = plan_alloc(c, p, v.size());
    if (rc == HFK_OK && !v.empty() &&
        hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = hfk_set_error(HFK_EHIP, "plan upload failed");
    return rc;
}
template <typename T>
static int entry_alloc(hfk_ctx* c, PlanEntry& e, T** p, size_t n)
{
    (void)e;
    return plan_alloc(c, p, n);
}

void hfk_plan_free_arrays(hfk_ctx* c)
{
    if (c->plan_chunks.empty()) return;
    hipStreamSynchronize(c->stream);
    for (void* q : c->plan_chunks) hipFree(q);
    c->plan_chunks.clear();
    c->plan_cur = nullptr;
    c->plan_left = 0;
}

static void free_plan(hfk_ctx* c)
{
    for (PlanEntry& e : c->plan)
        for (void* p : e.owned) dev_free(c, p);
    c->plan.clear();
    hfk_plan_free_arrays(c);
    if (c->d_kpart) dev_free(c, c->d_kpart);
    if (c->d_jpart) dev_free(c, c->d_jpart);
    c->d_kpart = c->d_jpart = nullptr;
    c->slot_bytes = 0;
}

// The launch plan: every class pair (bra class >= ket class); with the
// mixed-precision threshold set, its tasks form an FP64 and an FP32 part
// (two launches over one set of slots).
int hfk_ctx_plan(hfk_ctx* c)
{
    if (const char* v = getenv("HFK_KETS_PER_LANE")) c->kets_per_lane = std::max(1, atoi(v));
    free_plan(c);
    const bool ref = c->opt.screen_mode == HFK_SCREEN_REFERENCE && c->opt.screen_tau > 0.0;
    for (int kb = 0; kb < HFK_NPC; ++kb) {
        if (c->class_count[kb] == 0) continue;
        for (int kk = 0; kk <= kb; ++kk) {
            if (c->class_count[kk] == 0) continue;
            int la, lb, lc, ld;
            hfk_pair_class_l(kb, &la, &lb);
            hfk_pair_class_l(kk, &lc, &ld);
            const bool coop = hfk_class_coop(la, lb, lc, ld);   // FP64 only
            PlanEntry e = {};
            e.kb = kb;
            e.kk = kk;
            e.la = la;
            e.lb = lb;
            e.lc = lc;
            e.ld = ld;
            e.split = c->opt.fp32_threshold > 0.0 && !coop;   // FP64 and FP32 task parts
            const long long nb = c->class_count[kb], nk = c->class_count[kk];
            e.nquart = (kb == kk) ? nb * (nb + 1) / 2 : nb * nk;
            e.mode = ref ? 1 : 0;   // 1: kernels apply the reference screen
            c->plan.push_back(e);
        }
    }
    const int nshell = (int)c->shells.size();
    // Task lists of the J/K (and gradient) kernels, built on the device
    // (hfk_plan.hip).  Class lists are ordered by primitive-pair count, then
    // by descending screen value, so for each bra pair the kets surviving the
    // pair screen are a prefix of each segment of equal primitive-pair count.
    // A task is a bra pair, a list of runs (such prefixes, as position ranges
    // in the ket class list) and primitive ranges of bra and ket.  Lanes walk
    // the runs in order, so a wave mostly meets one primitive-pair count.
    //
    // Slots.  Each task stores its K' rows (the bra shells' rows over the
    // launch's compact columns) and its J'_ab; each (bra, ket, primitive
    // piece) stores its J'_cd in an entry laid out by bra-list order: bra i
    // and ket segment s own the entries jbase[i][s] + piece * len + (p -
    // first) for its kets p in [first, jend[i][s]).  That layout lets the
    // J'_cd reduction walk, for a ket position, the bras that cover it
    // without any per-quartet index.
    for (PlanEntry& e : c->plan) {
        const bool coop = hfk_class_coop(e.la, e.lb, e.lc, e.ld);
        const int block = hfk_plan_block(c, e);
        const int NA = ncart(e.la), NB = ncart(e.lb), NC = ncart(e.lc), ND = ncart(e.ld);
        const int nrow = NA + NB, nbra_comp = NA * NB, nket_comp = NC * ND;
        // compact K columns: the functions of the ket class's shells
        std::vector<int> colmap(c->n + 1, -1), collist;
        for (int S = 0; S < nshell; ++S) {
            const HostShell& h = c->shells[S];
            if (h.L != e.lc && h.L != e.ld) continue;
            for (int x = 0; x < HFK_MAXCOMP; ++x)
                if (h.bf[x] >= 0) colmap[h.bf[x]] = 1;
        }
        for (int f = 0; f < c->n; ++f)
            if (colmap[f] > 0) {
                colmap[f] = (int)collist.size();
                collist.push_back(f);
            }
        e.ncols = (int)collist.size();
        for (int f = 0; f <= c->n; ++f)
            if (colmap[f] < 0) colmap[f] = e.ncols;   // never a ket column (or the padding): discarded
        HFK_TRY(entry_upload(c, e, &e.d_colmap, colmap));
        HFK_TRY(entry_upload(c, e, &e.d_collist, collist));
        e.d_kets = c->d_class[e.kk];   // runs index the ket class list itself (not owned)
        HFK_TRY(hfk_plan_entry_device(c, e));
        // LDS of one workgroup: the K' rows must fit; the bra shells' density
        // rows join them while the workgroups of four waves per SIMD still fit
        // in the 160 KB of a CU
        {
            const size_t lds = coop ? hfk_coop_jk_lds_bytes(e.la, e.lb, e.ncols) + 16 * 1024
                                    : hfk_jk_lds_bytes(e.la, e.lb, e.ncols, e.bra_stage, 0);
            if (lds > 64 * 1024 || e.ncols > 4096)
                return hfk_set_error(HFK_EUNSUP,
                                     "(%d%d|%d%d): K rows of %d columns need %zu bytes of LDS per workgroup (at most 64 KB)",
                                     e.la, e.lb, e.lc, e.ld, e.ncols, lds);
            e.drows = !coop && hfk_jk_lds_bytes(e.la, e.lb, e.ncols, e.bra_stage, 1) <= hfk_lds_budget(block) &&
                      !getenv("HFK_NO_DROWS");
        }
        const long long nk = c->class_count[e.kk];
        HFK_TRY(entry_alloc(c, e, &e.jl.part, std::max<size_t>(1, (size_t)e.jl.nchunk * nk * nket_comp)));
        // the slots
        HFK_TRY(entry_alloc(c, e, &e.d_kslot, std::max<size_t>(1, (size_t)e.ntask * nrow * e.ncols)));
        HFK_TRY(entry_alloc(c, e, &e.d_jab, std::max<size_t>(1, (size_t)e.ntask * nbra_comp)));
        HFK_TRY(entry_alloc(c, e, &e.d_jcd, std::max<size_t>(1, (size_t)e.njcd * nket_comp)));
        e.jl.jcd = e.d_jcd;
        e.slot_bytes = sizeof(double) * ((size_t)e.ntask * (nrow * e.ncols + nbra_comp) + (size_t)e.nkjob * e.ncols +
                                         ((size_t)e.njcd + (size_t)e.jl.nchunk * nk) * nket_comp);
        c->slot_bytes += e.slot_bytes;
    }
    // the launches' partial J', K'
    {
        const size_t np1 = (size_t)(c->n + 1) * (c->n + 1), nl = std::max<size_t>(1, c->plan.size());
        HFK_TRY(dev_alloc(&c->d_kpart, nl * np1));
        HFK_TRY(dev_alloc(&c->d_jpart, nl * np1));
        // every build writes the same elements of the partials; the rest stay zero
        hfk_launch_fill(c->d_kpart, 0.0, (long long)(nl * np1), c->stream);
        hfk_launch_fill(c->d_jpart, 0.0, (long long)(nl * np1), c->stream);
        HFK_HIP(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < c->plan.size(); ++k) {
            PlanEntry& e = c->plan[k];
            JKPart& P = e.jkp;
            memset(&P, 0, sizeof P);
            P.kslot = e.d_kslot;
            P.csr_off = e.d_kcsr_off;
            P.csr = e.d_kcsr;
            P.kjob = e.d_kjob;
            P.nkjob = e.nkjob;
            P.kseg = e.d_kseg;
            P.krow = e.d_krow;
            P.nkrow = e.nkrow;
            P.collist = e.d_collist;
            P.ncols = e.ncols;
            P.nrow = ncart(e.la) + ncart(e.lb);
            P.nshell = nshell;
            P.jab = e.d_jab;
            P.bcsr_off = e.d_bcsr_off;
            P.jpart = e.jl.part;
            P.nchunk = e.jl.nchunk;
            P.bra_list = c->d_class[e.kb];
            P.ket_list = c->d_class[e.kk];
            P.nbra = (int)c->class_count[e.kb];
            P.nket = (int)c->class_count[e.kk];
            P.same = e.kb == e.kk;
            P.na = ncart(e.la);
            P.nb = ncart(e.lb);
            P.nc = ncart(e.lc);
            P.nd = ncart(e.ld);
            P.Kpart = c->d_kpart + k * np1;
            P.Jpart = c->d_jpart + k * np1;
        }
    }
    c->counted = false;   // exact counts on first request (hfk_ctx_count)
    c->first_launch = -1;
    c->serial_runs = 0;
    const size_t np = c->plan.size();
    // event pool for per-launch timing
    while (c->events.size() < 6 * np + 2) {   // J/K parts, then the gradient launches
        hipEvent_t ev;
        HFK_HIP(hipEventCreate(&ev));
        c->events.push_back(ev);
    }
    return HFK_OK;
}

void hfk_plan_args(hfk_ctx* c, const PlanEntry& e, QuartetArgs& a)
{
    a.tasks = e.d_tasks;
    a.runs = e.d_runs;
    a.ket_idx = e.d_kets;
    a.ntask = e.ntask;
    a.task0 = 0;
    a.bra_stage = e.bra_stage;
    a.task_len_max = e.task_len_max;
    a.block = hfk_plan_block(c, e);
    a.kslot = e.d_kslot;
    a.jab = e.d_jab;
    a.jab_slot = e.d_jab_slot;
    a.jcd = e.d_jcd;
    a.colmap = e.d_colmap;
    a.ncols = e.ncols;
    a.ksoa = c->d_ksoa[e.kk];
    a.kdesc = e.d_kdesc;
    a.kpos_q = e.d_kpos_q;
    a.kscale = e.d_kscale;
    a.collist = e.d_collist;
    a.drows = e.drows;
    a.bra_list = c->d_class[e.kb];
    a.ket_list = c->d_class[e.kk];
    a.nbra = c->class_count[e.kb];
    a.nket = c->class_count[e.kk];
    a.same = (e.kb == e.kk);
    a.nquart = e.nquart;
    a.prec_sel = 0;   // the precision is the task part's (planned), not chosen per quartet
    a.fp32_thr = c->opt.fp32_threshold;
}

// Tasks of one precision part of a plan entry: [first, first + count).
static void part_range(const PlanEntry& e, int part, long long& first, long long& count)
{
    first = part == 0 ? 0 : e.ntask64;
    count = part == 0 ? e.ntask64 : e.ntask - e.ntask64;
}

// Run the plan for one operation (J/K or flat export); per-launch HIP events.
int hfk_ctx_run_quartets(hfk_ctx* c, QuartetArgs a, int op)
{
    if (op == HFK_OP_JK && !a.D) return hfk_set_error(HFK_EINVAL, "J/K launch without a density");
    long long total = 0;
    const size_t np = c->plan.size();
    // J/K launches write only their own slots, so they may run concurrently.
    // The longest class (as measured in the context's second J/K run; the
    // first two are serial and the first pays one-time costs) runs first,
    // alone on the context stream: it fills the GPU by itself and keeps a
    // clean time.  The others then run round-robin on the auxiliary streams,
    // forked from and joined back into the context stream, and fill each
    // other's tails.  A class's FP64 and FP32 parts run one after the other
    // on its stream, followed by the fixed-order reductions of its slots.
    const bool conc = op == HFK_OP_JK && c->concurrent && np > 2 && c->first_launch >= 0;
    const size_t first = conc ? (size_t)c->first_launch : 0;
    int slot = 0, missing = -1, lds_fail = -1;
    for (size_t oi = 0; oi < np && missing < 0 && lds_fail < 0; ++oi) {
        // launch order: the first class, then the others in plan order
        const size_t i = !conc ? oi : oi == 0 ? first : oi <= first ? oi - 1 : oi;
        const PlanEntry& e = c->plan[i];
        if (conc && oi == 1) {
            HFK_HIP(hipEventRecord(c->ev_fork, c->stream));
            for (int k = 0; k < hfk_ctx::NAUX; ++k) HFK_HIP(hipStreamWaitEvent(c->aux[k], c->ev_fork, 0));
        }
        hfk_plan_args(c, e, a);
        total += e.nquart;
        hipStream_t s = (conc && oi > 0) ? c->aux[slot++ % hfk_ctx::NAUX] : c->stream;
        int rc = 0;
        // exports (flat array, Schwarz diagonal) walk the pair lists once, in FP64
        const int nparts = (op == HFK_OP_JK && e.split) ? 2 : 1;
        for (int part = 0; part < nparts && rc == 0; ++part) {
            QuartetArgs b = a;
            if (op == HFK_OP_JK) {
                long long t0, cnt;
                part_range(e, part, t0, cnt);
                b.tasks = e.d_tasks + t0;
                b.task0 = t0;
                b.ntask = cnt;
                // the FP32 part of an s/p class runs at more waves per SIMD
                // than the density rows in LDS would leave room for
                if (part == 1 && e.la <= 1) b.drows = 0;
            }
            HFK_HIP(hipEventRecord(c->events[4 * i + 2 * part], s));
            rc = hfk_launch_quartets(e.la, e.lb, e.lc, e.ld, op, part, c->opt.boys_mode,
                                     op == HFK_OP_JK ? e.mode : 0, b, s);
            HFK_HIP(hipEventRecord(c->events[4 * i + 2 * part + 1], s));
        }
        if (rc == -3) lds_fail = (int)i;
        else if (rc) missing = (int)i;
        c->plan[i].solo = (!conc || oi == 0) ? 1 : 0;
        if (op == HFK_OP_JK && rc == 0) {
            hfk_launch_reduce_jcd(e.jl, s);
            hfk_launch_jk_part(e.jkp, c->n, c->d_sh_bf, c->d_sh_bfp, c->d_pair_A, c->d_pair_B, s);
        }
    }
    if (conc)
        for (int k = 0; k < hfk_ctx::NAUX; ++k) {
            HFK_HIP(hipEventRecord(c->ev_join[k], c->aux[k]));
            HFK_HIP(hipStreamWaitEvent(c->stream, c->ev_join[k], 0));
        }
    if (lds_fail >= 0) {
        const PlanEntry& e = c->plan[lds_fail];
        return hfk_set_error(HFK_EUNSUP, "(%d%d|%d%d): the K rows of %d columns do not fit in LDS", e.la, e.lb, e.lc,
                             e.ld, e.ncols);
    }
    if (missing >= 0) {
        const PlanEntry& e = c->plan[missing];
        return hfk_set_error(HFK_EUNSUP, "no kernel for class (%d%d|%d%d)", e.la, e.lb, e.lc, e.ld);
    }
    HFK_HIP(hipGetLastError());
    c->stats.quartets_launched = total;
    c->timed_op = op;
    c->timed_serial = !conc;
    return HFK_OK;
}

// After a synchronised run: per-launch milliseconds into the plan.
int hfk_ctx_collect_times(hfk_ctx* c)
{
    for (size_t i = 0; i < c->plan.size(); ++i) {
        PlanEntry& e = c->plan[i];
        for (int part = 0; part < 2; ++part) {
            e.ms[part] = 0.0;
            if (part == 1 && !(c->timed_op == HFK_OP_JK && e.split)) continue;
            float ms = 0.f;
            HFK_HIP(hipEventElapsedTime(&ms, c->events[4 * i + 2 * part], c->events[4 * i + 2 * part + 1]));
            e.ms[part] = ms;
        }
    }
    // the second serial J/K run ranks the classes for the concurrent runs that follow
    if (c->timed_op == HFK_OP_JK && c->timed_serial && c->first_launch < 0 && !c->plan.empty() &&
        ++c->serial_runs >= 2) {
        size_t first = 0;
        auto t = [&](size_t i) { return c->plan[i].ms[0] + c->plan[i].ms[1]; };
        for (size_t i = 1; i < c->plan.size(); ++i)
            if (t(i) > t(first)) first = i;
        c->first_launch = (int)first;
    }
    return HFK_OK;
}

// Exact counts (shell, primitive and canonical function quartets) per
// launch, over the task lists: only the surviving quartets are visited.  For
// statistics only, so computed on the first request rather than at planning.
static int hfk_ctx_count(hfk_ctx* c)
{
    if (c->counted) return HFK_OK;
    unsigned long long* d_cnt = nullptr;
    const size_t np = c->plan.size(), nslot = 6 * std::max<size_t>(np, 1);   // 3 counts x 2 parts
    HFK_TRY(dev_alloc(&d_cnt, nslot));
    hfk_launch_fill((double*)d_cnt, 0.0, (long long)nslot, c->stream);
    QuartetArgs a = hfk_ctx_args(c);
    a.sch_cut = c->opt.schwarz_cut;
    for (size_t i = 0; i < np; ++i) {
        PlanEntry& e = c->plan[i];
        hfk_plan_args(c, e, a);
        for (int part = 0; part < (e.split ? 2 : 1); ++part) {
            QuartetArgs b = a;
            long long t0, cnt;
            part_range(e, part, t0, cnt);
            b.tasks = e.d_tasks + t0;
            b.task0 = t0;
            b.ntask = cnt;
            hfk_launch_count_tasks(b, ncart(e.la), ncart(e.lb), ncart(e.lc), ncart(e.ld), d_cnt + 6 * i + 3 * part,
                                   c->stream);
        }
    }
    HFK_HIP(hipGetLastError());
    std::vector<unsigned long long> h(nslot);
    HFK_HIP(hipMemcpyAsync(h.data(), d_cnt, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    dev_free(c, d_cnt);
    long long fq = 0;
    for (size_t i = 0; i < np; ++i)
        for (int part = 0; part < 2; ++part) {
            PlanEntry& e = c->plan[i];
            e.shell_q[part] = (long long)h[6 * i + 3 * part];
            e.prim_q[part] = (long long)h[6 * i + 3 * part + 1];
            e.func_q[part] = (long long)h[6 * i + 3 * part + 2];
            fq += e.func_q[part];
        }
    c->stats.func_quartets = fq;
    c->counted = true;
    return HFK_OK;
}

// One record per launch: each plan entry's FP64 part, then its FP32 part
// when the entry is split.
extern "C" int hfk_get_class_stats(hfk_ctx* c, hfk_class_stat* out, int32_t cap, int32_t* n)
{
    if (!c || !n) return hfk_set_error(HFK_EINVAL, "hfk_get_class_stats: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HFK_TRY(hfk_ctx_count(c));
    int32_t k = 0;
    for (size_t i = 0; i < c->plan.size(); ++i) {
        const PlanEntry& e = c->plan[i];
        for (int part = 0; part < (e.split ? 2 : 1); ++part, ++k) {
            if (!out || k >= cap) continue;
            hfk_class_stat& s = out[k];
            long long t0, cnt;
            part_range(e, part, t0, cnt);
            s.la = e.la;
            s.lb = e.lb;
            s.lc = e.lc;
            s.ld = e.ld;
            s.prec = part;
            s.solo = e.solo;
            s.shell_quartets = e.shell_q[part];
            s.prim_quartets = e.prim_q[part];
            s.func_quartets = e.func_q[part];
            s.ms = e.ms[part];
            s.tasks = cnt;
            s.grad_ms = part == 0 ? e.grad_ms : 0.0;
        }
    }
    *n = k;
    return HFK_OK;
}

// Device Fock build: G (or H+G) flat from a flat density, all on device.
int hfk_ctx_fock_device(hfk_ctx* c, const double* H_dev, const double* D_dev_flat, double* F_dev_flat, double extra_cut)
{
    // timed from here: the set-up launch (padded D), the class launches, each
    // followed on its stream by the fixed-order reductions of its slots into
    // its partial J', K', and the sum of the partials into G
    HFK_HIP(hipEventRecord(c->ev0, c->stream));
    hfk_launch_fock_prepare(c->n, D_dev_flat, c->d_D, c->stream);
    QuartetArgs a = hfk_ctx_args(c);
    a.sch_cut = std::max(c->opt.schwarz_cut, extra_cut);
    a.D = c->d_D;
    HFK_TRY(hfk_ctx_run_quartets(c, a, HFK_OP_JK));
    hfk_launch_symmetrize_parts(c->n, c->d_jpart, c->d_kpart, (int)c->plan.size(), H_dev, F_dev_flat, c->stream);
    HFK_HIP(hipEventRecord(c->ev1, c->stream));
    HFK_HIP(hipGetLastError());
    HFK_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HFK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.last_eri_ms = ms;
    HFK_TRY(hfk_ctx_collect_times(c));
    return HFK_OK;
}

extern "C" int hfk_fock_dev(hfk_ctx* c, const double* H_dev, const double* D_dev, double* F_dev)
{
    if (!c || !D_dev || !F_dev) return hfk_set_error(HFK_EINVAL, "hfk_fock_dev: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    auto t0 = std::chrono::steady_clock::now();
    int rc = hfk_ctx_fock_device(c, H_dev, D_dev, F_dev);
    c->stats.last_fock_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

extern "C" int hfk_fock(hfk_ctx* c, const double* H_flat, const double* D_flat, double* F_flat)
{
    if (!c || !D_flat || !F_flat) return hfk_set_error(HFK_EINVAL, "hfk_fock: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    auto t0 = std::chrono::steady_clock::now();
    const size_t nf = (size_t)c->n * (c->n + 1) / 2;
    HFK_HIP(hipMemcpyAsync(c->d_flatA, D_flat, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
    const double* Hd = nullptr;
    if (H_flat) {
        HFK_HIP(hipMemcpyAsync(c->d_flatC, H_flat, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
        Hd = c->d_flatC;
    }
    HFK_TRY(hfk_ctx_fock_device(c, Hd, c->d_flatA, c->d_flatB));
    HFK_HIP(hipMemcpyAsync(F_flat, c->d_flatB, sizeof(double) * nf, hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    c->stats.last_fock_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HFK_OK;
}

extern "C" int hfk_fock_G(hfk_ctx* c, const double* D_flat, double* G_flat)
{
    return hfk_fock(c, nullptr, D_flat, G_flat);
}

extern "C" int hfk_eri_flat(hfk_ctx* c, double* out, size_t out_len)
{
    if (!c || !out) return hfk_set_error(HFK_EINVAL, "hfk_eri_flat: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return hfk_ctx_eri_flat(c, out, out_len);
}

// calcIntegrals into a host array (the caller holds the context lock).
int hfk_ctx_eri_flat(hfk_ctx* c, double* out, size_t out_len)
{
    const long long n = c->n, m = n * (n + 1) / 2, len = m * (m + 1) / 2;
    if ((long long)out_len != len)
        return hfk_set_error(HFK_EINVAL, "hfk_eri_flat: out_len %zu, need %lld", out_len, len);
    // flat_base[i][j] = rank of (i,j,i,j) in calcIntegrals order
    std::vector<long long> base((size_t)n * n, 0);
    long long r = 0;
    for (long long i = 0; i < n; ++i)
        for (long long j = i; j < n; ++j) {
            base[(size_t)i * n + j] = r;
            r += (n - 1 - i) * (n - i) / 2 + (n - j);
        }
    long long* d_base = nullptr;
    double* d_flat = nullptr;
    HFK_TRY(dev_upload(c, &d_base, base));
    HFK_TRY(dev_alloc(&d_flat, (size_t)len));
    hfk_launch_fill(d_flat, 0.0, len, c->stream);
    QuartetArgs a = hfk_ctx_args(c);
    a.flat = d_flat;
    a.flat_base = d_base;
    int rc = hfk_ctx_run_quartets(c, a, HFK_OP_FLAT);
    if (rc == HFK_OK) {
        if (hipMemcpyAsync(out, d_flat, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = hfk_set_error(HFK_EHIP, "hfk_eri_flat: copy-out failed");
    }
    dev_free(c, d_flat);
    dev_free(c, d_base);
    return rc;
}

extern "C" int hfk_get_stats(hfk_ctx* c, hfk_stats* st)
{
    if (!c || !st) return hfk_set_error(HFK_EINVAL, "hfk_get_stats: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HFK_TRY(hfk_ctx_count(c));
    *st = c->stats;
    return HFK_OK;
}

OneEArgs hfk_onee_args(hfk_ctx* c)
{
    OneEArgs a;
    a.n = c->n;
    a.center = c->d_center;
    a.lxyz = c->d_lxyz;
    a.prim_off = c->d_prim_off;
    a.coef = c->d_coef;
    a.expo = c->d_expo;
    a.natom = (int)c->Z.size();
    a.Z = c->d_Z;
    a.R = c->d_R;
    a.bt = c->bt;
    a.boys = c->opt.boys_mode;
    a.vpair = c->d_vpair;
    a.vpair_off = c->vpair_off.data();
    return a;
}

int hfk_ctx_overlap_dev(hfk_ctx* c, double* S_dev)
{
    hfk_launch_overlap(hfk_onee_args(c), S_dev, c->stream);
    HFK_HIP(hipGetLastError());
    return HFK_OK;
}

int hfk_ctx_hcore_dev(hfk_ctx* c, double* H_dev)
{
    if (c->Z.empty()) return hfk_set_error(HFK_EINVAL, "hfk_hcore: basis has no nuclei (natom/Z/R)");
    hfk_launch_hcore(hfk_onee_args(c), H_dev, c->stream);
    HFK_HIP(hipGetLastError());
    return HFK_OK;
}

extern "C" int hfk_overlap(hfk_ctx* c, double* S_flat)
{
    if (!c || !S_flat) return hfk_set_error(HFK_EINVAL, "hfk_overlap: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    const size_t nf = (size_t)c->n * (c->n + 1) / 2;
    HFK_TRY(hfk_ctx_overlap_dev(c, c->d_flatA));
    HFK_HIP(hipMemcpyAsync(S_flat, c->d_flatA, sizeof(double) * nf, hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    return HFK_OK;
}

extern "C" int hfk_hcore(hfk_ctx* c, double* H_flat)
{
    if (!c || !H_flat) return hfk_set_error(HFK_EINVAL, "hfk_hcore: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    const size_t nf = (size_t)c->n * (c->n + 1) / 2;
    HFK_TRY(hfk_ctx_hcore_dev(c, c->d_flatA));
    HFK_HIP(hipMemcpyAsync(H_flat, c->d_flatA, sizeof(double) * nf, hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    return HFK_OK;
}
