// Setup kernels: Boys tables, primitive-pair records with the reference's
// per-component screen values, and small elementwise helpers.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "hfk_internal.hpp"
#include "hfk_quartet_common.hpp"

namespace {

// ------------------------------------------------------------------
// Boys tables
// ------------------------------------------------------------------

// Reference node: boysF m x (NumericalTools/Boys.hs:54-66) with kummer
// by 256-point Gauss-Legendre (Boys.hs:106-118, 155-157, 163-165).  glx/glw are the
// 128 positive nodes/weights in ascending order.
__device__ double boys_ref_node(int m, double x, const double* glx, const double* glw)
{
#pragma clang fp contract(off)
    if (m == 0) {
        if (x > 50.0) return 0.5 * sqrt(HFK_PI / x);   // asymptBoysF 0 x
        if (x > 0.0) return 0.5 * sqrt(HFK_PI / x) * erf(sqrt(x));
        return 1.0;
    }
    const double a = m + 0.5, b = m + 1.5, z = -x;
    const double gammaFactor = tgamma(b) / (tgamma(a) * tgamma(b - a));
    const double e = exp(1.0);
    double s = 0.0;
    for (int i = 0; i < 128; ++i) {
        const double xp = 0.5 * ((1.0 * glx[i]) + 1.0 + 0.0);
        const double xm = 0.5 * ((1.0 * -glx[i]) + 1.0 + 0.0);
        const double fp = pow(e, z * xp) * pow(1.0 - xp, b - a - 1.0) * pow(xp, a - 1.0);
        const double fm = pow(e, z * xm) * pow(1.0 - xm, b - a - 1.0) * pow(xm, a - 1.0);
        s += glw[i] * (fp + fm);
    }
    const double integral = 0.5 * (1.0 - 0.0) * s;
    return (gammaFactor * integral) / ((2.0 * m) + 1.0);
}

// Accurate F_m(x), x <= 50: e^-x sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)).
__device__ double boys_acc_node(int m, double x)
{
    double term = 1.0 / (2.0 * m + 1.0), s = term;
    for (int k = 1; k < 1000; ++k) {
        term *= 2.0 * x / (2.0 * m + 2.0 * k + 1.0);
        s += term;
        if (term < 1e-18 * s) break;
    }
    return exp(-x) * s;
}

__global__ void k_boys_tables(const double* glx, const double* glw, double* ref64, float* ref32, double* acc64,
                              float* acc32)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nref = BOYS_REF_M * BOYS_NX;
    if (t < nref) {
        const int m = t / BOYS_NX, i = t % BOYS_NX;
        const double v = boys_ref_node(m, 0.1 * (double)i, glx, glw);
        ref64[(size_t)i * BOYS_REF_STRIDE + m] = v;
        ref32[(size_t)i * BOYS_REF_STRIDE + m] = (float)v;
    } else if (t < nref + BOYS_ACC_M * BOYS_ACC_NX) {
        const int u = t - nref;
        const int m = u / BOYS_ACC_NX, i = u % BOYS_ACC_NX;
        const double v = boys_acc_node(m, 0.1 * (double)i);
        acc64[(size_t)i * BOYS_ACC_STRIDE + m] = v;
        acc32[(size_t)i * BOYS_ACC_STRIDE + m] = (float)v;
    }
}

// ------------------------------------------------------------------
// Primitive pairs
// ------------------------------------------------------------------

__device__ __forceinline__ int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ void cart_d(int l, int c, int* e)
{
    // canonical order: x^l first, decreasing x power, y before z
    int k = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (k == c) { e[0] = i; e[1] = j; e[2] = l - i - j; return; }
            ++k;
        }
}

// recursiveHermCoeff (Integrals/IntegralsEvaluation.hs:457-477), tabulated
// bottom-up for one axis: E[i][j][t], i<=2, j<=2, t<=4.
__device__ void herm_axis(double gamma, double xpa, double xpb, double e0, int li, int lj, double E[3][3][5])
{
#pragma clang fp contract(off)
    for (int i = 0; i <= li; ++i)
        for (int j = 0; j <= lj; ++j)
            for (int t = 0; t <= 4; ++t) {
                double v;
                if (i == 0 && j == 0 && t == 0) v = e0;
                else if (t > i + j) v = 0.0;
                else if (t == 0 && i >= 1) v = xpa * E[i - 1][j][0] + E[i - 1][j][1];
                else if (t == 0 && j >= 1) v = xpb * E[i][j - 1][0] + E[i][j - 1][1];
                else {
                    const double av = (i >= 1) ? E[i - 1][j][t - 1] : 0.0;
                    const double bv = (j >= 1) ? E[i][j - 1][t - 1] : 0.0;
                    v = (1.0 / (2.0 * gamma * (double)t)) * ((double)i * av + (double)j * bv);
                }
                E[i][j][t] = v;
            }
}

// twoTermsERI (IntegralsEvaluation.hs:319-337): the reference's screen value
// for primitives (ca,ea) with exponents la on A and (cb,eb) with lb on B.
__device__ double two_terms(const double* A, const int* la, double ca, double ea, const double* B, const int* lb,
                            double cb, double eb)
{
#pragma clang fp contract(off)
    const double p = ea + eb;
    const double mu = ea * eb * (1.0 / (ea + eb));
    double E[3][3][3][5];
    for (int x = 0; x < 3; ++x) {
        const double rp = (ea * A[x] + eb * B[x]) / (ea + eb);
        const double d = A[x] - B[x];
        const double e0 = exp(-(mu * (d * d)));
        herm_axis(p, rp - A[x], rp - B[x], e0, la[x], lb[x], E[x]);
    }
    double suma = 0.0;
    for (int t = 0; t <= la[0] + lb[0]; ++t)
        for (int u = 0; u <= la[1] + lb[1]; ++u)
            for (int v = 0; v <= la[2] + lb[2]; ++v) {
                const double c = ((1.0 * E[0][la[0]][lb[0]][t]) * E[1][la[1]][lb[1]][u]) * E[2][la[2]][lb[2]][v];
                suma += (((t + u + v) & 1) ? -1.0 : 1.0) * c;
            }
    double acc = 0.0;
    for (int t = 0; t <= la[0] + lb[0]; ++t)
        for (int u = 0; u <= la[1] + lb[1]; ++u)
            for (int v = 0; v <= la[2] + lb[2]; ++v) {
                const double c = ((1.0 * E[0][la[0]][lb[0]][t]) * E[1][la[1]][lb[1]][u]) * E[2][la[2]][lb[2]][v];
                acc += c * suma;
            }
    const double r = HFK_PI / (2.0 * p);
    const double r2 = r * r;
    const double r5 = (r2 * r2) * r;   // GHC (^5): x^5 = (x^2)^2 * x
    const double cte = 2.0 * ((ca * ca) * (cb * cb)) * sqrt(r5);
    return cte * acc;
}

__global__ void k_pair_setup(ShellDev sh, int npair, const int* pair_A, const int* pair_B, const int* pair_pp0,
                             const int* pair_q0, double* pp, double* qcomp, double* pair_qhi)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const int A = pair_A[p], B = pair_B[p];
    const int LA = sh.L[A], LB = sh.L[B];
    const int na = ncart_d(LA), nb = ncart_d(LB);
    const double* RA = sh.center + 3 * A;
    const double* RB = sh.center + 3 * B;
    double qmax_pair = 0.0;
    int k = 0;
    for (int ia = 0; ia < sh.nprim[A]; ++ia) {
        const double ea = sh.expo[sh.prim0[A] + ia], ca = sh.coef[sh.prim0[A] + ia];
        for (int ib = 0; ib < sh.nprim[B]; ++ib, ++k) {
            const double eb = sh.expo[sh.prim0[B] + ib], cb = sh.coef[sh.prim0[B] + ib];
            double* r = pp + (size_t)(pair_pp0[p] + k) * PP_STRIDE;
            const double pe = ea + eb;
            const double mu = ea * eb / pe;
            double R2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                const double P = (ea * RA[x] + eb * RB[x]) / pe;
                r[PP_PX + x] = P;
                r[PP_AX + x] = P - RA[x];
                r[PP_BX + x] = P - RB[x];
                const double d = RA[x] - RB[x];
                R2 += d * d;
            }
            r[PP_P] = pe;
            r[PP_K] = ca * cb * exp(-mu * R2) / pe;
            r[PP_OO2P] = 1.0 / (2.0 * pe);
            const int qoff = pair_q0[p] + k * na * nb;
            r[PP_QOFF] = (double)qoff;
            r[PP_EA] = ea;
            double qlo = INFINITY, qhi = 0.0;
            for (int a = 0; a < na; ++a) {
                int ea3[3];
                cart_d(LA, a, ea3);
                const double cac = ca * sh.scale[A * HFK_MAXCOMP + a];
                for (int b = 0; b < nb; ++b) {
                    int eb3[3];
                    cart_d(LB, b, eb3);
                    const double cbc = cb * sh.scale[B * HFK_MAXCOMP + b];
                    const double q = sqrt(two_terms(RA, ea3, cac, ea, RB, eb3, cbc, eb));
                    qcomp[qoff + a * nb + b] = q;
                    const double qq = isnan(q) ? INFINITY : q;
                    qlo = fmin(qlo, qq);
                    qhi = fmax(qhi, qq);
                }
            }
            r[PP_QLO] = qlo;
            r[PP_QHI] = qhi;
            qmax_pair = fmax(qmax_pair, qhi);
        }
    }
    pair_qhi[p] = qmax_pair;
}

__global__ void k_fill(double* p, double v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// G = (J' + J'^T) - 1/2 (K' + K'^T); G_flat (upper triangle, row-major) =
// H_flat + G if H is given; F_full (optional) receives the symmetric matrix.
__global__ void k_symmetrize_g(int n, const double* J, const double* K, const double* H, double* G_flat,
                               double* F_full)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        // row i: starts at i*n - i(i-1)/2
        const double nn = (double)n + 0.5;
        long long i = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
        if (i < 0) i = 0;
        while (i > 0 && i * n - i * (i - 1) / 2 > idx) --i;
        while ((i + 1) * n - (i + 1) * i / 2 <= idx) ++i;
        const long long j = i + (idx - (i * n - i * (i - 1) / 2));
        const long long ld = n + 1;
        const double jv = J[i * ld + j] + J[j * ld + i];
        const double kv = K[i * ld + j] + K[j * ld + i];
        double g = jv - 0.5 * kv;
        if (H) g += H[idx];
        G_flat[idx] = g;
        if (F_full) {
            F_full[i * n + j] = g;
            F_full[j * n + i] = g;
        }
    }
}

__global__ void k_flat_to_full(int n, const double* flat, double* full)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        const double nn = (double)n + 0.5;
        long long i = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
        if (i < 0) i = 0;
        while (i > 0 && i * n - i * (i - 1) / 2 > idx) --i;
        while ((i + 1) * n - (i + 1) * i / 2 <= idx) ++i;
        const long long j = i + (idx - (i * n - i * (i - 1) / 2));
        full[i * n + j] = flat[idx];
        full[j * n + i] = flat[idx];
    }
}

// One element idx < n(n+1)/2 + n + 1 of the flat -> padded (n+1)^2 conversion.
__device__ __forceinline__ void flat_to_padded_elem(int n, const double* flat, double* full, long long idx)
{
    const long long m = (long long)n * (n + 1) / 2;
    const long long ld = n + 1;
    if (idx >= m) {   // zero row/column n
        const long long t = idx - m;
        full[n * ld + t] = 0.0;
        full[t * ld + n] = 0.0;
        return;
    }
    const double nn = (double)n + 0.5;
    long long i = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
    if (i < 0) i = 0;
    while (i > 0 && i * n - i * (i - 1) / 2 > idx) --i;
    while ((i + 1) * n - (i + 1) * i / 2 <= idx) ++i;
    const long long j = i + (idx - (i * n - i * (i - 1) / 2));
    full[i * ld + j] = flat[idx];
    full[j * ld + i] = flat[idx];
}

__global__ void k_flat_to_padded(int n, const double* flat, double* full)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m + n + 1;
         idx += (long long)gridDim.x * blockDim.x)
        flat_to_padded_elem(n, flat, full, idx);
}

// Start of a Fock build in one launch: padded D from the flat density, and
// J', K' ((n+1)^2 each) and the ket-side sums (jk_len) cleared.
__global__ void k_fock_prepare(int n, const double* D_flat, double* D, double* J, double* K, double* Jk, long long jk_len)
{
    const long long m = (long long)n * (n + 1) / 2, np1 = (long long)(n + 1) * (n + 1);
    const long long total = max(max(m + n + 1, np1), jk_len);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        if (idx < m + n + 1) flat_to_padded_elem(n, D_flat, D, idx);
        if (idx < np1) {
            J[idx] = 0.0;
            K[idx] = 0.0;
        }
        if (idx < jk_len) Jk[idx] = 0.0;
    }
}

__global__ void k_full_to_flat(int n, const double* full, double* flat)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        const double nn = (double)n + 0.5;
        long long i = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
        if (i < 0) i = 0;
        while (i > 0 && i * n - i * (i - 1) / 2 > idx) --i;
        while ((i + 1) * n - (i + 1) * i / 2 <= idx) ++i;
        const long long j = i + (idx - (i * n - i * (i - 1) / 2));
        flat[idx] = full[i * n + j];
    }
}

// Counts over the task lists: one 64-lane workgroup per task (grid-stride).
__global__ void __launch_bounds__(64) k_count_tasks(QuartetArgs a, int na, int nb, int nc, int nd,
                                                    unsigned long long* out)
{
    unsigned long long cs = 0, cp = 0, cf = 0;
    for (long long t = blockIdx.x; t < a.ntask; t += gridDim.x) {
        const int4 task = a.tasks[t];
        const int bp = task.x;
        const int b0 = (task.w >> 24) & 0xff, b1 = (task.w >> 16) & 0xff;
        const int k0 = (task.w >> 8) & 0xff, k1 = task.w & 0xff;
        const bool first = (b0 == 0 && k0 == 0);   // each quartet counted once
        const double* bra = a.pp + (size_t)a.pair_pp0[bp] * PP_STRIDE;
        const int len = task_ket_count(a, task);
        for (int j = (int)threadIdx.x; j < len; j += blockDim.x) {
            const int kp = a.ket_idx[task_ket_pos(a, task, j)];
            if (!quartet_selected(a, bp, kp)) continue;
            const double* ket = a.pp + (size_t)a.pair_pp0[kp] * PP_STRIDE;
            const int ke = task_kprim_end(a, k1, kp);
            for (int x = b0; x < b1; ++x) {
                const double bq = bra[(size_t)x * PP_STRIDE + PP_QHI];
                for (int y = k0; y < ke; ++y)
                    if (!(a.qtau > 0.0 && bq * ket[(size_t)y * PP_STRIDE + PP_QHI] < a.qtau)) cp += 1;
            }
            if (first) {
                cs += 1;
                const unsigned long long nbra = (a.pair_A[bp] == a.pair_B[bp]) ? (unsigned long long)na * (na + 1) / 2
                                                                                  : (unsigned long long)na * nb;
                const unsigned long long nket = (a.pair_A[kp] == a.pair_B[kp]) ? (unsigned long long)nc * (nc + 1) / 2
                                                                                  : (unsigned long long)nc * nd;
                cf += (bp == kp) ? nbra * (nbra + 1) / 2 : nbra * nket;
            }
        }
    }
    if (cs) atomicAdd(out + 0, cs);
    if (cp) atomicAdd(out + 1, cp);
    if (cf) atomicAdd(out + 2, cf);
}

unsigned grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

void hfk_launch_pair_setup(const ShellDev& sh, int npair, const int* pair_A, const int* pair_B,
                           const int* pair_pp0, const int* pair_q0, double* pp, double* qcomp,
                           double* pair_qhi, hipStream_t s)
{
    if (npair <= 0) return;
    hipLaunchKernelGGL(k_pair_setup, dim3((npair + 127) / 128), dim3(128), 0, s, sh, npair, pair_A, pair_B,
                       pair_pp0, pair_q0, pp, qcomp, pair_qhi);
}

void hfk_launch_boys_tables(const double* glx, const double* glw, double* ref64, float* ref32, double* acc64,
                            float* acc32, hipStream_t s)
{
    const int total = BOYS_REF_M * BOYS_NX + BOYS_ACC_M * BOYS_ACC_NX;
    hipLaunchKernelGGL(k_boys_tables, dim3((total + 127) / 128), dim3(128), 0, s, glx, glw, ref64, ref32, acc64,
                       acc32);
}

void hfk_launch_count_tasks(const QuartetArgs& a, int na, int nb, int nc, int nd, unsigned long long* out,
                            hipStream_t s)
{
    if (a.ntask <= 0) return;
    const long long g = std::min<long long>(a.ntask, 8192);
    hipLaunchKernelGGL(k_count_tasks, dim3((unsigned)g), dim3(64), 0, s, a, na, nb, nc, nd, out);
}

void hfk_launch_fill(double* p, double v, long long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void hfk_launch_symmetrize_g(int n, const double* J, const double* K, const double* H, double* G_flat,
                             double* F_full, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_symmetrize_g, dim3(grid_for(m, 256)), dim3(256), 0, s, n, J, K, H, G_flat, F_full);
}

// Ket-side J'_cd sums, kept by position in the ket class list during the
// J/K launch, added into J' (atomic: absent components of different pairs
// share the zero row/column n).
// grid.y runs over the launches
__global__ void k_scatter_jk(double* J, const double* Jk, int copies, const JkScatter* ent, const int* pair_A,
                             const int* pair_B, const int* sh_bfp, int ld)
{
    const JkScatter e = ent[blockIdx.y];
    const int nz = e.nc * e.nd;
    const long long count = e.nket * nz;
    const double* src = Jk + e.off;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int c = 0; c < copies; ++c) v += src[(size_t)c * count + i];
        if (v == 0.0) continue;
        const long long pos = i / nz;
        const int z = (int)(i - pos * nz);
        const int kp = e.ket_list[pos];
        const int ic = sh_bfp[pair_A[kp] * HFK_MAXCOMP + z / e.nd], id = sh_bfp[pair_B[kp] * HFK_MAXCOMP + z % e.nd];
        atomic_add_f64(&J[(size_t)ic * ld + id], v);
    }
}

void hfk_launch_scatter_jk(double* J, const double* Jk, int copies, const JkScatter* ent, int nent, long long max_count,
                           const int* pair_A, const int* pair_B, const int* sh_bfp, int ld, hipStream_t s)
{
    if (nent <= 0 || max_count <= 0) return;
    hipLaunchKernelGGL(k_scatter_jk, dim3(grid_for(max_count, 256), (unsigned)nent), dim3(256), 0, s, J, Jk, copies, ent,
                       pair_A, pair_B, sh_bfp, ld);
}

void hfk_launch_flat_to_full(int n, const double* flat, double* full, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_flat_to_full, dim3(grid_for(m, 256)), dim3(256), 0, s, n, flat, full);
}

void hfk_launch_flat_to_padded(int n, const double* flat, double* full, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2 + n + 1;
    hipLaunchKernelGGL(k_flat_to_padded, dim3(grid_for(m, 256)), dim3(256), 0, s, n, flat, full);
}

void hfk_launch_fock_prepare(int n, const double* D_flat, double* D, double* J, double* K, double* Jk, long long jk_len,
                             hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2, np1 = (long long)(n + 1) * (n + 1);
    const long long total = std::max(std::max(m + n + 1, np1), jk_len);
    hipLaunchKernelGGL(k_fock_prepare, dim3(grid_for(total, 256)), dim3(256), 0, s, n, D_flat, D, J, K, Jk, jk_len);
}

void hfk_launch_full_to_flat(int n, const double* full, double* flat, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_full_to_flat, dim3(grid_for(m, 256)), dim3(256), 0, s, n, full, flat);
}
