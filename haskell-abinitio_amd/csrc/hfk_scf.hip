// SCF-side entry points: density, symmetric orthogonalisation, unitary
// transform and the DIIS-accelerated RHF driver (HartreeFock/HartreeFock.hs:
// 70-135,195-287, HartreeFock/DIIS.hs:55-137), all FP64 on the device.  The
// eigen-solve is rocSOLVER dsyevd; the dense products are the MFMA GEMM.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <deque>
#include <mutex>
#include <vector>

#include "hfk.h"
#include "hfk_host.hpp"

// rocBLAS handles of destroyed contexts, per device (a handle and its device
// workspace are costly to create; short-lived contexts reuse them).  A
// handle is owned by one context at a time and bound to its stream.
namespace {
constexpr size_t kBlasPoolMax = 16;
std::mutex g_blas_mu;
std::vector<std::pair<int, rocblas_handle>> g_blas_pool;
}  // namespace

int hfk_linalg_init(hfk_ctx* c)
{
    if (c->blas) return HFK_OK;
    rocblas_handle h = nullptr;
    {
        std::lock_guard<std::mutex> g(g_blas_mu);
        for (size_t i = g_blas_pool.size(); i-- > 0;)
            if (g_blas_pool[i].first == c->device) {
                h = g_blas_pool[i].second;
                g_blas_pool.erase(g_blas_pool.begin() + (long)i);
                break;
            }
    }
    if (!h && rocblas_create_handle(&h) != rocblas_status_success)
        return hfk_set_error(HFK_EHIP, "rocblas_create_handle failed");
    rocblas_set_stream(h, c->stream);
    c->blas = (void*)h;
    return HFK_OK;
}

// Called with the context's stream synchronised.
void hfk_linalg_destroy(hfk_ctx* c)
{
    rocblas_handle h = (rocblas_handle)c->blas;
    c->blas = nullptr;
    if (!h) return;
    rocblas_set_stream(h, nullptr);
    {
        std::lock_guard<std::mutex> g(g_blas_mu);
        if (g_blas_pool.size() < kBlasPoolMax) {
            g_blas_pool.emplace_back(c->device, h);
            return;
        }
    }
    rocblas_destroy_handle(h);
}

// Symmetric eigen-solve (EigenValues.hs:29-31 + VectorTools.hs:27-31):
// eigenvalues ascending; on return row k of A is the k-th eigenvector.
// E (n doubles) and info (two ints) are device workspace.  With info_host
// (pinned) the status is copied there asynchronously and the caller checks
// it after its next stream synchronisation; without, this call waits.
// Up to n = 32 (molecular fragments) rocSOLVER's Jacobi solver is the
// faster one (0.15 against 0.33 ms at n = 13), above it divide and conquer
// (1.6 against 2.8 ms for Jacobi-D&C at n = 104); both sort ascending.
int hfk_dev_syev(hfk_ctx* c, int n, double* A, double* w, double* E, int* info, int* info_host)
{
    HFK_TRY(hfk_linalg_init(c));
    // A is symmetric, so row-major == column-major on input; the solvers leave
    // eigenvectors in columns (column-major) == rows in our row-major view.
    rocblas_status st;
    if (n <= 32)
        st = rocsolver_dsyevj((rocblas_handle)c->blas, rocblas_esort_ascending, rocblas_evect_original,
                              rocblas_fill_upper, n, A, n, 0.0 /* machine precision */, E /* residual */, 100,
                              info + 1 /* sweeps */, w, info);
    else
        st = rocsolver_dsyevd((rocblas_handle)c->blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, E, info);
    if (st != rocblas_status_success) return hfk_set_error(HFK_EHIP, "rocsolver eigen-solve failed (%d)", (int)st);
    if (info_host) {
        HFK_HIP(hipMemcpyAsync(info_host, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        return HFK_OK;
    }
    int hinfo = 0;
    HFK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    if (hinfo != 0) return hfk_set_error(HFK_ENUMERIC, "rocsolver eigen-solve info=%d", hinfo);
    return HFK_OK;
}

namespace {

struct Work {
    hfk_ctx* c;
    int n;
    double *S = nullptr, *X = nullptr, *H = nullptr;          // S, X full; H flat
    double *T1 = nullptr, *T2 = nullptr, *T3 = nullptr, *Ev = nullptr, *w = nullptr;
    double *Ffull = nullptr, *C = nullptr;
    double *part = nullptr, *scal = nullptr;
    double* ev_e = nullptr;    // eigen-solve workspace (n)
    int* ev_info = nullptr;
    // pinned host mirror of scal (the context's buffer), and the status of
    // the last eigen-solve, copied asynchronously and checked after the next
    // synchronisation
    double* hscal = nullptr;
    int* hinfo = nullptr;
};

constexpr int SCAL_LEN = 64;
constexpr int SCAL_DOTS = 8;   // scal[SCAL_DOTS + k]: trace(e_cur^T e_k) for DIIS slot k

int work_alloc(hfk_ctx* c, Work& W)
{
    W.c = c;
    W.n = c->n;
    const size_t nn = (size_t)c->n * c->n, nf = (size_t)c->n * (c->n + 1) / 2;
    HFK_TRY(dev_alloc(&W.S, nn));
    HFK_TRY(dev_alloc(&W.X, nn));
    HFK_TRY(dev_alloc(&W.H, nf));
    HFK_TRY(dev_alloc(&W.T1, nn));
    HFK_TRY(dev_alloc(&W.T2, nn));
    HFK_TRY(dev_alloc(&W.T3, nn));
    HFK_TRY(dev_alloc(&W.Ev, nn));
    HFK_TRY(dev_alloc(&W.w, (size_t)c->n));
    HFK_TRY(dev_alloc(&W.Ffull, nn));
    HFK_TRY(dev_alloc(&W.C, nn));
    HFK_TRY(dev_alloc(&W.part, 1024));
    HFK_TRY(dev_alloc(&W.scal, SCAL_LEN));
    HFK_TRY(dev_alloc(&W.ev_e, (size_t)c->n));
    HFK_TRY(dev_alloc(&W.ev_info, 2));
    static_assert(SCAL_LEN <= HFK_PIN_DOUBLES, "pinned scalar buffer");
    W.hscal = c->h_pin;   // the context's (calls on a context are serialised)
    W.hinfo = (int*)(c->h_pin + HFK_PIN_DOUBLES);
    *W.hinfo = 0;
    return HFK_OK;
}

void work_free(Work& W)
{
    hfk_ctx* c = W.c;
    hipStreamSynchronize(c->stream);
    double* ps[] = {W.S, W.X, W.H, W.T1, W.T2, W.T3, W.Ev, W.w, W.Ffull, W.C, W.part, W.scal, W.ev_e};
    for (double* p : ps) dev_free(c, p);
    dev_free(c, W.ev_info);
    W.hscal = nullptr;
    W.hinfo = nullptr;
}

// Wait for the stream and copy scal[first, first+count) to W.hscal; also
// checks the status of the eigen-solves queued since the last call.
int sync_read(Work& W, int first, int count)
{
    hfk_ctx* c = W.c;
    HFK_HIP(hipMemcpyAsync(W.hscal + first, W.scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    if (*W.hinfo != 0) return hfk_set_error(HFK_ENUMERIC, "rocsolver eigen-solve info=%d", *W.hinfo);
    return HFK_OK;
}

void gemm(hfk_ctx* c, int m, int n, int k, double alpha, const double* A, int lda, int ta, const double* B, int ldb,
          int tb, double beta, double* C, int ldc)
{
    hfk_launch_dgemm(m, n, k, alpha, A, lda, ta, B, ldb, tb, beta, C, ldc, c->stream);
}

// X = U s^-1/2 U^T  (BasisOrthogonalization.hs:26-35); S full (n x n) in W.S
int symm_ortho_dev(Work& W, const double* S, double* X)
{
    hfk_ctx* c = W.c;
    const int n = W.n;
    HFK_HIP(hipMemcpyAsync(W.Ev, S, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c->stream));
    HFK_TRY(hfk_dev_syev(c, n, W.Ev, W.w, W.ev_e, W.ev_info, nullptr));
    // rows of Ev are eigenvectors u_k; X = sum_k u_k s_k^-1/2 u_k^T = Ev^T diag(s^-1/2) Ev
    HFK_HIP(hipMemcpyAsync(W.T1, W.Ev, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c->stream));
    hfk_launch_scale_cols(n, W.T1, W.w, 0 | 4, c->stream);   // scale row k by s_k^-1/2
    gemm(c, n, n, n, 1.0, W.Ev, n, 1, W.T1, n, 0, 0.0, X, n);
    HFK_HIP(hipGetLastError());
    return HFK_OK;
}

// Fp = X^T F X  (LinearAlgebra.hs:155-159); F full
void unitary_dev(Work& W, const double* X, const double* F, double* Fp)
{
    const int n = W.n;
    gemm(W.c, n, n, n, 1.0, F, n, 0, X, n, 0, 0.0, W.T3, n);       // F X
    gemm(W.c, n, n, n, 1.0, X, n, 1, W.T3, n, 0, 0.0, Fp, n);      // X^T (F X)
}

// D = 2 C_occ C_occ^T  (HartreeFock.hs:195-208), D full
void density_dev(Work& W, const double* C, int occ, double* Dfull)
{
    const int n = W.n;
    gemm(W.c, n, n, occ, 2.0, C, n, 0, C, n, 1, 0.0, Dfull, n);
}

// diagonalHF (HartreeFock.hs:263-273): F flat -> C (W.C), eps (W.w), D flat
int diagonal_hf(Work& W, const double* Fflat, int occ, double* Dflat)
{
    hfk_ctx* c = W.c;
    const int n = W.n;
    hfk_launch_flat_to_full(n, Fflat, W.Ffull, c->stream);
    unitary_dev(W, W.X, W.Ffull, W.Ev);
    HFK_TRY(hfk_dev_syev(c, n, W.Ev, W.w, W.ev_e, W.ev_info, W.hinfo));
    gemm(c, n, n, n, 1.0, W.X, n, 0, W.Ev, n, 1, 0.0, W.C, n);   // C = X C', C' columns = eigenvectors
    density_dev(W, W.C, occ, W.T1);
    hfk_launch_full_to_flat(n, W.T1, Dflat, c->stream);
    HFK_HIP(hipGetLastError());
    return HFK_OK;
}

// gaussElem (NumericalTools/GaussElimination.hs:42-112): elimination without
// pivoting; fails when a row has no entry above 1e-6 or a NaN/Inf appears.
bool gauss_elem(std::vector<double> A, std::vector<double> b, int m, std::vector<double>& x)
{
    for (int k = 0; k < m; ++k) {
        std::vector<double> A2 = A, b2 = b;
        for (int i = k + 1; i < m; ++i) {
            const double alpha = A[(size_t)i * m + k] / A[(size_t)k * m + k];
            for (int j = k; j < m; ++j) A2[(size_t)i * m + j] = A[(size_t)i * m + j] - alpha * A[(size_t)k * m + j];
            b2[i] = b[i] - alpha * b[k];
        }
        A = A2;
        b = b2;
    }
    for (int i = 0; i < m; ++i) {
        bool any = false;
        for (int j = 0; j < m; ++j) any = any || fabs(A[(size_t)i * m + j]) > 1e-6;
        if (!any) return false;
    }
    for (double v : A)
        if (std::isnan(v) || std::isinf(v)) return false;
    x.assign(m, 0.0);
    for (int i = m - 1; i >= 0; --i) {
        double s = 0.0;
        for (int j = i + 1; j < m; ++j) s += x[j] * A[(size_t)i * m + j];
        x[i] = (b[i] - s) * (1.0 / A[(size_t)i * m + i]);
    }
    return true;
}

}  // namespace

extern "C" void hfk_default_scf_options(hfk_scf_options* o)
{
    o->guess = HFK_GUESS_CORE;
    o->max_steps = 30;      // HartreeFock.hs:87
    o->diis_size = 5;       // HartreeFock.hs:78
    o->diis_start = 0.1;    // DIIS.hs:58
    o->verbose = 0;
    o->pad = 0;
    o->guess_density = nullptr;
    o->incremental_tau = 0.0;   // the reference rebuilds F from D every step
    o->rebuild_every = 16;
    o->pad2 = 0;
    o->energy_converge = 0.0;   // the reference tests max|e| only
    o->converge = 1.0e-7;   // HartreeFock.hs:113
    o->charge = 0.0;
}

// After a synchronised call: the GEMM time between the context's ev0 and
// ev1 and its flop count into the statistics.
static int record_gemm_time(hfk_ctx* c, double flops)
{
    float ms = 0.f;
    HFK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.last_gemm_ms = ms;
    c->stats.last_gemm_flops = flops;
    return HFK_OK;
}

extern "C" int hfk_density(hfk_ctx* c, const double* C, int32_t occ, double* D_flat)
{
    if (!c || !C || !D_flat || occ < 0 || occ > c->n) return hfk_set_error(HFK_EINVAL, "hfk_density: bad argument");
    std::lock_guard<std::mutex> g(c->mu);
    Work W;
    int rc = work_alloc(c, W);
    const int n = c->n;
    if (rc == HFK_OK) {
        hipMemcpyAsync(W.C, C, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream);
        hipEventRecord(c->ev0, c->stream);
        density_dev(W, W.C, occ, W.T1);
        hipEventRecord(c->ev1, c->stream);
        hfk_launch_full_to_flat(n, W.T1, c->d_flatA, c->stream);
        hipMemcpyAsync(D_flat, c->d_flatA, sizeof(double) * n * (n + 1) / 2, hipMemcpyDeviceToHost, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = hfk_set_error(HFK_EHIP, "hfk_density: HIP failure");
        else rc = record_gemm_time(c, 2.0 * n * n * (double)occ);
    }
    work_free(W);
    return rc;
}

extern "C" int hfk_symm_ortho(hfk_ctx* c, const double* S, double* X)
{
    if (!c || !S || !X) return hfk_set_error(HFK_EINVAL, "hfk_symm_ortho: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    Work W;
    int rc = work_alloc(c, W);
    const int n = c->n;
    if (rc == HFK_OK) {
        hipMemcpyAsync(W.S, S, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream);
        rc = symm_ortho_dev(W, W.S, W.X);
        if (rc == HFK_OK) {
            hipMemcpyAsync(X, W.X, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream);
            if (hipStreamSynchronize(c->stream) != hipSuccess) rc = hfk_set_error(HFK_EHIP, "hfk_symm_ortho: HIP failure");
        }
    }
    work_free(W);
    return rc;
}

extern "C" int hfk_unitary(hfk_ctx* c, const double* X, const double* F_flat, double* Fp)
{
    if (!c || !X || !F_flat || !Fp) return hfk_set_error(HFK_EINVAL, "hfk_unitary: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    Work W;
    int rc = work_alloc(c, W);
    const int n = c->n;
    if (rc == HFK_OK) {
        hipMemcpyAsync(W.X, X, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream);
        hipMemcpyAsync(c->d_flatA, F_flat, sizeof(double) * n * (n + 1) / 2, hipMemcpyHostToDevice, c->stream);
        hfk_launch_flat_to_full(n, c->d_flatA, W.Ffull, c->stream);
        hipEventRecord(c->ev0, c->stream);
        unitary_dev(W, W.X, W.Ffull, W.T1);
        hipEventRecord(c->ev1, c->stream);
        hipMemcpyAsync(Fp, W.T1, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = hfk_set_error(HFK_EHIP, "hfk_unitary: HIP failure");
        else rc = record_gemm_time(c, 4.0 * n * (double)n * n);
    }
    work_free(W);
    return rc;
}

// ---------------------------------------------------------------------------
// The reference's index maps exactly as written, including their behaviour
// for indices outside a matrix of the given dimension, which the reference's
// Harris guess produces (NumericalTools/LinearAlgebra.hs:194-283).
namespace refidx {
typedef long long ll;
// sum (takeZero w (iterate pred dim)): takeZero 0 = [0], take of a negative count = []
static ll calc_elem(ll dim, ll w)
{
    ll s = 0;
    for (ll t = 0; t < w; ++t) s += dim - t;
    return s;
}
static ll dim2_to_flat(ll dim, ll x, ll y)   // indexDIM2toFlat
{
    return x <= y ? calc_elem(dim, x) + (y - x) : calc_elem(dim, y) + (x - y);
}
static bool flat_to_dim2(ll dim, ll idx, ll& x, ll& y)   // indexFlat2DIM2; false if it would not end
{
    ll r = idx, k = 0;
    while (!(r < dim - k)) {
        if (dim - k <= 0) return false;
        r -= dim - k;
        ++k;
    }
    x = k;
    y = idx + x - calc_elem(dim, x);
    return true;
}
static ll sum_index_triang(ll n) { return (n * n + n) / 2; }   // n(n+1) >= 0: div = floor
static ll sum_first_index(ll n)
{
    ll s = n + sum_index_triang(n - 1);
    for (ll k = 1; k <= n - 1; ++k) s += sum_index_triang(n) - k;
    return s;
}
static ll four_index_to_flat(ll n, const ll q[4])   // fourIndex2Flat
{
    const ll i = q[0], j = q[1], k = q[2], l = q[3];
    ll a0 = 0;
    for (ll m = 0; m <= i - 1; ++m) a0 += sum_first_index(n - m);   // totalFirstIndex
    ll pos;
    if (i == j && i == k) {
        pos = l - i;
    } else if (i == j) {
        pos = (n - i) + dim2_to_flat(n - (i + 1), k - (i + 1), l - (i + 1));
    } else {
        ll s = 0;
        for (ll m = 1; m <= (j - 1) - i; ++m) s += sum_index_triang(n - i) - m;
        pos = (n - i) + sum_index_triang((n - 1) - i) + s + (dim2_to_flat(n, k, l) - dim2_to_flat(n, i, j));
    }
    return a0 + pos;
}
static void sort_keys(ll i, ll j, ll k, ll l, ll q[4])   // sortKeys (HartreeFock.hs:153-156)
{
    const ll a0 = std::min(i, j), a1 = std::max(i, j), b0 = std::min(k, l), b1 = std::max(k, l);
    if (a0 < b0 || (a0 == b0 && a1 <= b1)) { q[0] = a0; q[1] = a1; q[2] = b0; q[3] = b1; }
    else { q[0] = b0; q[1] = b1; q[2] = a0; q[3] = a1; }
}
}   // namespace refidx

// harrisFunctional / computeGuessIntegrals as written (HartreeFock.hs:127-149):
// G_idx = 2 sum_{k<occ} [eri(i,j,k,k) - 1/2 eri(i,k,k,j)] with (i,j) =
// indexFlat2DIM2 M idx (M, the flat length, as the dimension) and the flat
// positions taken by fourIndex2Flat with occ as the basis size.  The reference
// indexes with `!`, so an index outside the integral array aborts it; here that
// is HFK_EINVAL.  Host arithmetic on the flat ERI array (small n only).
static int reference_harris_g(hfk_ctx* c, int occ, std::vector<double>& G)
{
    const long long n = c->n, M = n * (n + 1) / 2, len = M * (M + 1) / 2;
    if (len > (1LL << 26))
        return hfk_set_error(HFK_EINVAL, "hfk_scf: HFK_GUESS_REFERENCE reads the flat ERI array (%lld entries); "
                                         "use it for small systems only", len);
    std::vector<double> eri((size_t)len);
    HFK_TRY(hfk_ctx_eri_flat(c, eri.data(), eri.size()));
    G.assign((size_t)M, 0.0);
    for (long long idx = 0; idx < M; ++idx) {
        long long i, j;
        if (!refidx::flat_to_dim2(M, idx, i, j))
            return hfk_set_error(HFK_EINVAL, "hfk_scf: reference Harris guess: indexFlat2DIM2 does not end");
        double s = 0.0;
        for (long long k = 0; k < occ; ++k) {
            long long q1[4], q2[4];
            refidx::sort_keys(i, j, k, k, q1);
            refidx::sort_keys(i, k, k, j, q2);
            const long long f1 = refidx::four_index_to_flat(occ, q1), f2 = refidx::four_index_to_flat(occ, q2);
            if (f1 < 0 || f1 >= len || f2 < 0 || f2 >= len)
                return hfk_set_error(HFK_EINVAL, "hfk_scf: reference Harris guess indexes the integral array at %lld / %lld "
                                                 "(length %lld); the reference aborts here", f1, f2, len);
            s += eri[(size_t)f1] - 0.5 * eri[(size_t)f2];
        }
        G[(size_t)idx] = 2.0 * s;
    }
    return HFK_OK;
}

// scfHF + scfDIIS (HartreeFock.hs:70-123)
extern "C" int hfk_scf(hfk_ctx* c, const hfk_scf_options* o, hfk_scf_result* res, double* C_out, double* eps_out,
                       double* D_out, double* F_out)
{
    if (!c || !res) return hfk_set_error(HFK_EINVAL, "hfk_scf: null argument");
    memset(res, 0, sizeof *res);
    if (c->Z.empty()) return hfk_set_error(HFK_EINVAL, "hfk_scf: basis has no nuclei");
    hfk_scf_options opt;
    if (o) opt = *o;
    else hfk_default_scf_options(&opt);
    std::lock_guard<std::mutex> g(c->mu);
    auto t0 = std::chrono::steady_clock::now();
    double fock_s = 0.0;
    const int n = c->n;
    const size_t nn = (size_t)n * n, nf = (size_t)n * (n + 1) / 2;
    double zsum = 0.0;
    for (double z : c->Z) zsum += z;
    const int occ = (int)floor((zsum - opt.charge) / 2.0);   // HartreeFock.hs:75-76
    if (occ <= 0 || occ > n) return hfk_set_error(HFK_EINVAL, "hfk_scf: occupied=%d for n=%d", occ, n);
    if (opt.guess != HFK_GUESS_CORE && opt.guess != HFK_GUESS_INTENDED && opt.guess != HFK_GUESS_DENSITY &&
        opt.guess != HFK_GUESS_REFERENCE)
        return hfk_set_error(HFK_EINVAL, "hfk_scf: unknown guess %d", opt.guess);
    if (opt.guess == HFK_GUESS_DENSITY && !opt.guess_density)
        return hfk_set_error(HFK_EINVAL, "hfk_scf: HFK_GUESS_DENSITY needs guess_density");
    if (!(opt.incremental_tau >= 0.0)) return hfk_set_error(HFK_EINVAL, "hfk_scf: incremental_tau must be >= 0");
    if (opt.diis_size < 1 || opt.diis_size > SCAL_LEN - SCAL_DOTS)
        return hfk_set_error(HFK_EINVAL, "hfk_scf: diis_size %d outside [1, %d]", opt.diis_size, SCAL_LEN - SCAL_DOTS);
    // nuclearRep (HartreeFock.hs:300-305)
    double vnn = 0.0;
    for (int a = 0; a < c->natom; ++a)
        for (int b = a + 1; b < c->natom; ++b) {
            double r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                const double d = c->R[3 * a + x] - c->R[3 * b + x];
                r2 += d * d;
            }
            vnn += c->Z[a] * c->Z[b] / sqrt(r2);
        }

    Work W;
    int rc = work_alloc(c, W);
    double *Dflat = nullptr, *Fflat = nullptr, *Fd = nullptr, *Dfull = nullptr, *Sfull = nullptr;
    double *Dprev = nullptr, *dD = nullptr;   // last D a Fock matrix was built from; D - Dprev
    std::vector<double*> errs, focks;   // DIIS history, newest first
    if (rc == HFK_OK) rc = dev_alloc(&Dflat, nf);
    if (rc == HFK_OK) rc = dev_alloc(&Fflat, nf);
    if (rc == HFK_OK) rc = dev_alloc(&Fd, nf);
    if (rc == HFK_OK) rc = dev_alloc(&Dfull, nn);
    if (rc == HFK_OK) rc = dev_alloc(&Sfull, nn);
    const bool incremental = opt.incremental_tau > 0.0;
    if (rc == HFK_OK) rc = dev_alloc(&Dprev, nf);
    if (rc == HFK_OK && incremental) rc = dev_alloc(&dD, nf);
    c->scf_cached = false;
    for (int k = 0; k < opt.diis_size && rc == HFK_OK; ++k) {
        double *e = nullptr, *f = nullptr;
        rc = dev_alloc(&e, nn);
        if (rc == HFK_OK) rc = dev_alloc(&f, nf);
        errs.push_back(e);
        focks.push_back(f);
    }
    std::vector<double*> free_e(errs), free_f(focks);
    std::deque<int> hist;   // indices into errs/focks, newest first
    std::vector<double> Bdot((size_t)std::max(opt.diis_size, 1) * std::max(opt.diis_size, 1), 0.0);
    auto run = [&]() -> int {
        HFK_TRY(hfk_ctx_hcore_dev(c, W.H));
        HFK_TRY(hfk_ctx_overlap_dev(c, c->d_flatA));
        hfk_launch_flat_to_full(n, c->d_flatA, Sfull, c->stream);
        HFK_TRY(symm_ortho_dev(W, Sfull, W.X));
        // initial density
        if (opt.guess == HFK_GUESS_CORE) {
            HFK_TRY(diagonal_hf(W, W.H, occ, Dflat));
        } else if (opt.guess == HFK_GUESS_REFERENCE) {
            std::vector<double> G;
            HFK_TRY(reference_harris_g(c, occ, G));
            HFK_HIP(hipMemcpyAsync(Fflat, G.data(), sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
            HFK_HIP(hipStreamSynchronize(c->stream));   // G is pageable and local
            hfk_launch_axpby((long long)nf, 1.0, W.H, 1.0, Fflat, c->stream);   // fockGuess = fcore + guess
            HFK_TRY(diagonal_hf(W, Fflat, occ, Dflat));
        } else if (opt.guess == HFK_GUESS_DENSITY) {
            HFK_HIP(hipMemcpyAsync(Dflat, opt.guess_density, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
        } else {
            hfk_launch_guess_density(n, occ, Dflat, c->stream);
            auto f0 = std::chrono::steady_clock::now();
            HFK_TRY(hfk_ctx_fock_device(c, W.H, Dflat, Fflat));
            fock_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - f0).count();
            HFK_TRY(diagonal_hf(W, Fflat, occ, Dflat));
        }
        bool diis_on = false;
        int since_full = 0;
        for (int step = 0; step < opt.max_steps; ++step) {
            auto f0 = std::chrono::steady_clock::now();
            // fock (HartreeFock.hs:213-222).  Incremental builds (not in the
            // reference): G is linear in D, so F(D_k) = F(D_k-1) + G(D_k - D_k-1);
            // a quartet contributes at most 4 |(ab|cd)| max|dD| <= 4 Q_ab Q_cd max|dD|
            // to any element, so those below tau / max|dD| are skipped.  Fflat
            // still holds F(D_k-1) (DIIS works on copies).
            if (incremental && step > 0 && since_full + 1 < std::max(1, opt.rebuild_every)) {
                HFK_HIP(hipMemcpyAsync(dD, Dflat, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream));
                hfk_launch_axpby((long long)nf, -1.0, Dprev, 1.0, dD, c->stream);
                hfk_launch_reduce((long long)nf, dD, nullptr, 1, W.part, W.scal + 3, c->stream);
                HFK_TRY(sync_read(W, 3, 1));
                const double dmax = W.hscal[3];
                const double cut = dmax > 0.0 ? opt.incremental_tau / dmax : INFINITY;
                HFK_TRY(hfk_ctx_fock_device(c, Fflat, dD, Fflat, cut));
                ++since_full;
            } else {
                HFK_TRY(hfk_ctx_fock_device(c, W.H, Dflat, Fflat));
                since_full = 0;
            }
            HFK_HIP(hipMemcpyAsync(Dprev, Dflat, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream));
            fock_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - f0).count();
            // calcErrorMtx (DIIS.hs:119-132): X^T (F D S - S D F) X
            hfk_launch_flat_to_full(n, Fflat, W.Ffull, c->stream);
            hfk_launch_flat_to_full(n, Dflat, Dfull, c->stream);
            gemm(c, n, n, n, 1.0, Dfull, n, 0, Sfull, n, 0, 0.0, W.T1, n);     // D S
            gemm(c, n, n, n, 1.0, W.Ffull, n, 0, W.T1, n, 0, 0.0, W.T2, n);   // F D S
            hfk_launch_antisym(n, W.T2, W.T1, c->stream);                      // FDS - SDF
            gemm(c, n, n, n, 1.0, W.T1, n, 0, W.X, n, 0, 0.0, W.T2, n);
            // a storage slot for the current error/Fock pair (the history
            // holds at most diis_size-1 entries, so one is always free)
            int cur = -1;
            for (int k = 0; k < opt.diis_size; ++k) {
                bool used = false;
                for (int h : hist) used = used || (h == k);
                if (!used) { cur = k; break; }
            }
            if (cur < 0) return hfk_set_error(HFK_EINVAL, "hfk_scf: DIIS history overflow");
            double* ecur = errs[cur];
            gemm(c, n, n, n, 1.0, W.X, n, 1, W.T2, n, 0, 0.0, ecur, n);
            HFK_HIP(hipMemcpyAsync(focks[cur], Fflat, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream));
            hfk_launch_reduce((long long)nn, ecur, nullptr, 1, W.part, W.scal, c->stream);
            // the DIIS matrix entries of the current error with itself and the
            // stored ones (trace(e_i^T e_j)); entries between stored errors were
            // computed when the newer of the two was current.  One wait for all.
            hfk_launch_reduce((long long)nn, ecur, ecur, 0, W.part, W.scal + SCAL_DOTS + cur, c->stream);
            for (int h : hist) hfk_launch_reduce((long long)nn, ecur, errs[h], 0, W.part, W.scal + SCAL_DOTS + h, c->stream);
            HFK_TRY(sync_read(W, 0, SCAL_DOTS + opt.diis_size));
            const double maxe = W.hscal[0];
            Bdot[(size_t)cur * opt.diis_size + cur] = W.hscal[SCAL_DOTS + cur];
            for (int h : hist)
                Bdot[(size_t)cur * opt.diis_size + h] = Bdot[(size_t)h * opt.diis_size + cur] = W.hscal[SCAL_DOTS + h];
            if (std::isnan(maxe)) return hfk_set_error(HFK_ENUMERIC, "hfk_scf: NaN in the DIIS error at step %d", step);
            // diisDriver (DIIS.hs:55-65)
            const double* Fuse = Fflat;
            auto update = [&](std::deque<int>& hs) {
                // updateDIIS (DIIS.hs:111-114): keep at most diis_size-1 stored after adding
                hs.push_front(cur);
                if ((int)hs.size() >= opt.diis_size) hs.pop_back();
            };
            if (!diis_on) {
                if (step > 0 && maxe < opt.diis_start) {
                    update(hist);
                    diis_on = true;
                }
            } else {
                // diis (DIIS.hs:70-88) with the current pair in front of the history
                std::deque<int> use(hist);
                for (;;) {
                    const int m = (int)use.size() + 1;
                    std::vector<int> idx;
                    idx.push_back(cur);
                    for (int h : use) idx.push_back(h);
                    std::vector<double> B((size_t)(m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0);
                    for (int i = 0; i < m; ++i)
                        for (int j = 0; j <= i; ++j) {
                            const double v = Bdot[(size_t)idx[i] * opt.diis_size + idx[j]];   // trace(e_i^T e_j)
                            B[(size_t)i * (m + 1) + j] = v;
                            B[(size_t)j * (m + 1) + i] = v;
                        }
                    for (int i = 0; i < m; ++i) {
                        B[(size_t)i * (m + 1) + m] = -1.0;
                        B[(size_t)m * (m + 1) + i] = -1.0;
                    }
                    rhs[m] = -1.0;
                    std::vector<double> cs;
                    if (gauss_elem(B, rhs, m + 1, cs)) {
                        hfk_launch_fill(Fd, 0.0, (long long)nf, c->stream);
                        for (int i = 0; i < m; ++i) hfk_launch_axpby((long long)nf, cs[i], focks[idx[i]], 1.0, Fd, c->stream);
                        Fuse = Fd;
                        break;
                    }
                    if (use.empty()) break;   // nothing left to drop: keep the plain Fock
                    use.pop_back();           // initS: drop the oldest and retry
                    hist = use;
                }
                update(hist);
            }
            // diagonalHF + variationalE (HartreeFock.hs:111-112, 277-287)
            HFK_TRY(diagonal_hf(W, Fuse, occ, Dflat));
            hfk_launch_flat_to_full(n, Dflat, Dfull, c->stream);
            hfk_launch_flat_to_full(n, Fuse, W.T1, c->stream);
            hfk_launch_flat_to_full(n, W.H, W.T2, c->stream);
            hfk_launch_axpby((long long)nn, 1.0, W.T2, 1.0, W.T1, c->stream);   // H + F
            hfk_launch_reduce((long long)nn, Dfull, W.T1, 0, W.part, W.scal + 2, c->stream);
            HFK_TRY(sync_read(W, 2, 1));
            const double etot = 0.5 * W.hscal[2] + vnn;
            res->delta_energy = (step == 0) ? INFINITY : etot - res->energy;
            res->energy = etot;
            res->e_nuc = vnn;
            res->max_error = maxe;
            res->steps = step + 1;
            res->occupied = occ;
            if (opt.verbose) {
                fprintf(stderr, "hfk_scf: step %d E %.10f max|e| %.3e%s\n", step + 1, etot, maxe, diis_on ? " diis" : "");
                fflush(stderr);
            }
            const bool energy_ok = opt.energy_converge <= 0.0 || std::fabs(res->delta_energy) < opt.energy_converge;
            if (step != 0 && maxe < opt.converge && energy_ok) {
                if (F_out) HFK_HIP(hipMemcpyAsync(F_out, Fuse, sizeof(double) * nf, hipMemcpyDeviceToHost, c->stream));
                if (D_out) HFK_HIP(hipMemcpyAsync(D_out, Dflat, sizeof(double) * nf, hipMemcpyDeviceToHost, c->stream));
                if (C_out) HFK_HIP(hipMemcpyAsync(C_out, W.C, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
                if (eps_out) HFK_HIP(hipMemcpyAsync(eps_out, W.w, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
                // keep (D, F(D)) of the last build for hfk_gradient
                if (!c->d_scf_D) HFK_TRY(dev_alloc(&c->d_scf_D, nf));
                if (!c->d_scf_F) HFK_TRY(dev_alloc(&c->d_scf_F, nf));
                HFK_HIP(hipMemcpyAsync(c->d_scf_D, Dprev, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream));
                HFK_HIP(hipMemcpyAsync(c->d_scf_F, Fflat, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream));
                HFK_HIP(hipStreamSynchronize(c->stream));
                c->scf_cached = true;
                return HFK_OK;
            }
        }
        return hfk_set_error(HFK_ESCF, "SCF maxium steps exceeded");   // HartreeFock.hs:123
    };
    if (rc == HFK_OK) rc = run();
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->fock_seconds = fock_s;
    for (double* p : free_e) dev_free(c, p);
    for (double* p : free_f) dev_free(c, p);
    dev_free(c, Dflat);
    dev_free(c, Fflat);
    dev_free(c, Fd);
    dev_free(c, Dfull);
    dev_free(c, Sfull);
    dev_free(c, Dprev);
    dev_free(c, dD);
    work_free(W);
    return rc;
}
