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

// Start of a Fock build: padded D ((n+1)^2, row/column n zero) from the flat density.
__global__ void k_fock_prepare(int n, const double* D_flat, double* D)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m + n + 1;
         idx += (long long)gridDim.x * blockDim.x)
        flat_to_padded_elem(n, D_flat, D, idx);
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

// Position-major ket records (hfk_launch_ket_soa): one thread per position.
__global__ void k_ket_soa(const double* __restrict__ pp, const double* __restrict__ qcomp,
                          const int* __restrict__ pair_pp0, const int* __restrict__ klist, int nket,
                          const long long* __restrict__ pos_soa, const int* __restrict__ pos_len,
                          const int* __restrict__ pos_npp, int nf, int ncomp, double* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nket) return;
    const int npp = pos_npp[j];
    const long long len = pos_len[j];
    double* o = out + pos_soa[j] + j;
    const double* rec = pp + (size_t)pair_pp0[klist[j]] * PP_STRIDE;
    for (int ik = 0; ik < npp; ++ik) {
        const double* r = rec + (size_t)ik * PP_STRIDE;
        double* w = o + (long long)ik * nf * len;
        for (int f = 0; f < PP_STRIDE; ++f) w[f * len] = r[f];
        const double* q = qcomp + (long long)r[PP_QOFF];
        for (int z = 0; z < ncomp; ++z) w[(PP_STRIDE + z) * len] = q[z];
    }
}

void hfk_launch_ket_soa(const double* pp, const double* qcomp, const int* pair_pp0, const int* klist, int nket,
                        const long long* pos_soa, const int* pos_len, const int* pos_npp, int nf, int ncomp,
                        double* out, hipStream_t s)
{
    if (nket <= 0) return;
    hipLaunchKernelGGL(k_ket_soa, dim3((nket + 255) / 256), dim3(256), 0, s, pp, qcomp, pair_pp0, klist, nket, pos_soa,
                       pos_len, pos_npp, nf, ncomp, out);
}

// Each pair's primitive-pair records in descending order of their largest
// screen value QHI (stable: ties keep their index order), written to out,
// with their per-component screen values moved along (qout), so that a
// pair's value blocks stay contiguous in record order (the J/K kernel stages
// a task's blocks in one copy).  With that order the reference screen
// !(q_ab q_cd < tau) fails, for a ket primitive, on a suffix of the bra
// primitives, and the per-lane ERI loops stop at the first failure.
__global__ void k_pair_sort(int npair, const int* pair_pp0, const int* pair_npp, const int* pair_q0, long long nq_tot,
                            const double* pp, const double* qcomp, double* out, double* qout)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const size_t base = (size_t)pair_pp0[p];
    const int n = pair_npp[p];
    if (n <= 0) return;
    const long long q0 = pair_q0[p], q1 = p + 1 < npair ? pair_q0[p + 1] : nq_tot;
    const int nc = (int)((q1 - q0) / n);   // screen values per record
    for (int k = 0; k < n; ++k) {
        const double qk = pp[(base + k) * PP_STRIDE + PP_QHI];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double qj = pp[(base + j) * PP_STRIDE + PP_QHI];
            rank += (qj > qk || (qj == qk && j < k)) ? 1 : 0;
        }
        const double* src = pp + (base + k) * PP_STRIDE;
        double* dst = out + (base + rank) * PP_STRIDE;
#pragma unroll
        for (int f = 0; f < PP_STRIDE; ++f) dst[f] = src[f];
        const long long qsrc = (long long)src[PP_QOFF], qdst = q0 + (long long)rank * nc;
        for (int z = 0; z < nc; ++z) qout[qdst + z] = qcomp[qsrc + z];
        dst[PP_QOFF] = (double)qdst;
    }
}

void hfk_launch_pair_sort(int npair, const int* pair_pp0, const int* pair_npp, const int* pair_q0, long long nq_tot,
                          const double* pp, const double* qcomp, double* out, double* qout, hipStream_t s)
{
    if (npair <= 0) return;
    hipLaunchKernelGGL(k_pair_sort, dim3((npair + 127) / 128), dim3(128), 0, s, npair, pair_pp0, pair_npp, pair_q0,
                       nq_tot, pp, qcomp, out, qout);
}

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


// ----------------------------------------------------------------------
// Fixed-order reductions of the J/K slots.  Every sum below runs over its
// terms in an order fixed by the plan (bra-list order, piece order, task
// order, plan order), and every output element is written by one thread
// with a plain store, so a Fock build is bitwise reproducible.
// ----------------------------------------------------------------------

// J'_cd, pass 1: workgroup = (tile of ptile consecutive positions of one
// ket segment, chunk of consecutive bras of one bra segment); thread =
// (position, component).  The chunk's ket ends and entry offsets for the
// segment are loaded into LDS first; the workgroup then walks the chunk's
// bras in order and a thread adds the bras that cover its position.  The
// bras' ket ends do not increase along a bra segment (bras sorted by
// descending screen value), so the walk stops at the first bra that covers
// none of the tile; for a class paired with itself only bras at or after a
// position hold it.
#define HFK_JCD_STAGE 256
__global__ void __launch_bounds__(256) k_reduce_jcd_part(JcdLayout L)
{
    __shared__ int s_end[HFK_JCD_STAGE];
    __shared__ long long s_base[HFK_JCD_STAGE];
    const int2 tile = L.tiles[blockIdx.x];
    const int4 ch = L.chunks[blockIdx.y];
    const int ks = tile.x, t0 = tile.y;
    const int first = L.kseg_first[ks], kend = L.kseg_first[ks + 1];
    const int nkc = L.nket_comp;
    const int pl = threadIdx.x / nkc, k = threadIdx.x - pl * nkc;
    const int pos = t0 + pl;
    const bool valid = pl < L.ptile && pos < kend;
    const int off = pos - first;
    const int np = L.npiece[ch.x * L.nkseg + ks];
    // the chunk's bras [ch.y, ch.z), staged HFK_JCD_STAGE at a time
    int i0 = ch.y;
    if (L.same) i0 = max(i0, t0);
    double sum = 0.0;
    bool done = false;   // uniform
    for (int base = i0; base < ch.z && !done; base += HFK_JCD_STAGE) {
        const int cnt = min(HFK_JCD_STAGE, ch.z - base);
        __syncthreads();
        for (int b = threadIdx.x; b < cnt; b += blockDim.x) {
            const long long qi = (long long)(base + b) * L.nkseg + ks;
            s_end[b] = L.jend[qi];
            s_base[b] = L.jbase[qi];
        }
        __syncthreads();
        if (np == 1) {
            // one piece per quartet: the loads of 8 bras are issued before
            // their adds (in bra order; a bra that does not cover the
            // position adds nothing)
            constexpr int U = 8;
            for (int b0 = 0; b0 < cnt; b0 += U) {
                if (s_end[b0] <= t0) {   // uniform: no later bra covers the tile
                    done = true;
                    break;
                }
                double v[U];
                bool in[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int b = b0 + u;
                    in[u] = b < cnt && valid && pos < s_end[b] && (!L.same || pos <= base + b);
                    v[u] = in[u] ? L.jcd[(s_base[b] + off) * nkc + k] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (in[u]) sum += v[u];
            }
        } else {
            for (int b = 0; b < cnt; ++b) {
                const int end = s_end[b];
                if (end <= t0) {   // uniform: no later bra covers the tile
                    done = true;
                    break;
                }
                if (valid && pos < end && (!L.same || pos <= base + b)) {
                    const long long len = end - first;
                    const long long e0 = s_base[b] + off;
                    constexpr int U = 8;
                    for (int p0 = 0; p0 < np; p0 += U) {
                        double v[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) v[u] = p0 + u < np ? L.jcd[(e0 + (p0 + u) * len) * nkc + k] : 0.0;
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (p0 + u < np) sum += v[u];
                    }
                }
            }
        }
    }
    if (valid) L.part[((long long)blockIdx.y * L.nket + pos) * nkc + k] = sum;
}

void hfk_launch_reduce_jcd(const JcdLayout& L, hipStream_t s)
{
    const long long per = (long long)L.nket * L.nket_comp;
    if (per <= 0) return;
    if (L.nchunk > 0 && L.ntile > 0)
        hipLaunchKernelGGL(k_reduce_jcd_part, dim3((unsigned)L.ntile, (unsigned)L.nchunk), dim3(256), 0, s, L);
}

// The launch's partial J', K'.
// k_jk_part, blocks [0, ceil(nkjob / G)): first pass of the K rows.  A job
// is HFK_KSEG consecutive entries (task, row offset) of one shell component
// row; its slot rows are summed in entry order.  The threads form G = 256 /
// (64 ceil(ncols / 64)) groups, one job per group, a thread owning compact
// columns in registers; the loads of 8 entries are issued before their adds
// (in entry order), so a lane keeps 8 reads in flight.
// Blocks above: J' partial, thread = (position in the bra class list, bra
// component) (bra side, plus the ket side of the same pair when the classes
// are one), then for two classes (position in the ket class list, ket
// component); the ket side sums the J'_cd chunk sums in chunk order.
// k_krow_final: second pass, thread = (row, compact column): the row's job
// sums in job order, stored to the row of the partial K'.
__device__ __forceinline__ void k_part_kjobs(const JKPart& P, int blk)
{
    constexpr int NCOL = 16;   // columns per thread: ncols <= 4096 (the planner checks)
    constexpr int U = 8;       // entries in flight per lane
    const int tid = threadIdx.x;
    const int width = 64 * ((P.ncols + 63) / 64);
    const int G = width >= 256 ? 1 : 256 / width;
    const int gw = G == 1 ? 256 : width;
    const int g = tid / gw, lane = tid - g * gw;
    const int job = blk * G + g;
    if (g >= G || job >= P.nkjob) return;
    const int4 jb = P.kjob[job];
    const int x = jb.x % HFK_MAXCOMP;
    const int ncols = P.ncols, nc = (ncols - lane + gw - 1) / gw;   // columns of this lane
    double acc[NCOL];
#pragma unroll
    for (int c = 0; c < NCOL; ++c) acc[c] = 0.0;
    for (int q0 = jb.y; q0 < jb.z; q0 += U) {
        const double* src[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = min(q0 + u, jb.z - 1);   // past the end: a valid entry, not added
            const int2 en = P.csr[q];
            src[u] = P.kslot + ((size_t)en.x * P.nrow + en.y + x) * ncols + lane;
        }
        const int nu = min(U, jb.z - q0);
#pragma unroll
        for (int c = 0; c < NCOL; ++c) {
            if (c >= nc) break;
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[u][gw * c];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (u < nu) acc[c] += v[u];
        }
    }
    double* out = P.kseg + (size_t)job * ncols + lane;
#pragma unroll
    for (int c = 0; c < NCOL; ++c)
        if (c < nc) out[gw * c] = acc[c];
}

__device__ __forceinline__ double k_part_ket_sum(const JKPart& P, int pos, int c)
{
    constexpr int U = 8;
    const long long per = (long long)P.nket * P.nc * P.nd;
    const double* src = P.jpart + (long long)pos * P.nc * P.nd + c;
    double v = 0.0;
    for (int q0 = 0; q0 < P.nchunk; q0 += U) {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = q0 + u < P.nchunk ? src[(q0 + u) * per] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u < P.nchunk) v += t[u];
    }
    return v;
}

__global__ void __launch_bounds__(256) k_jk_part(JKPart P, int n, const int* sh_bf, const int* sh_bfp,
                                                 const int* pair_A, const int* pair_B)
{
    const int width = 64 * ((P.ncols + 63) / 64);
    const int G = width >= 256 ? 1 : 256 / width;
    const int nkblk = (P.nkjob + G - 1) / G;
    if ((int)blockIdx.x < nkblk) {
        k_part_kjobs(P, blockIdx.x);
        return;
    }
    const int nbc = P.na * P.nb, nkc = P.nc * P.nd;
    const long long tb = (long long)P.nbra * nbc;
    const long long total = tb + (P.same ? 0 : (long long)P.nket * nkc);
    const long long t = (long long)(blockIdx.x - nkblk) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int p, x, y;
    double v = 0.0;
    if (t < tb) {
        const int pos = (int)(t / nbc), c = (int)(t - (long long)pos * nbc);
        p = P.bra_list[pos];
        x = c / P.nb;
        y = c - x * P.nb;
        for (int q = P.bcsr_off[pos]; q < P.bcsr_off[pos + 1]; ++q) v += P.jab[(size_t)q * nbc + c];
        if (P.same) v += k_part_ket_sum(P, pos, c);
    } else {
        const long long u = t - tb;
        const int pos = (int)(u / nkc), c = (int)(u - (long long)pos * nkc);
        p = P.ket_list[pos];
        x = c / P.nd;
        y = c - x * P.nd;
        v = k_part_ket_sum(P, pos, c);
    }
    const int ia = sh_bfp[pair_A[p] * HFK_MAXCOMP + x], ib = sh_bfp[pair_B[p] * HFK_MAXCOMP + y];
    if (ia < n && ib < n) P.Jpart[(size_t)ia * (n + 1) + ib] = v;
}

__global__ void __launch_bounds__(256) k_krow_final(JKPart P, int n, const int* sh_bf)
{
    constexpr int U = 8;
    const long long nt = (long long)P.nkrow * P.ncols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
         t += (long long)gridDim.x * blockDim.x) {
        const int row = (int)(t / P.ncols), col = (int)(t - (long long)row * P.ncols);
        const int4 kr = P.krow[row];
        const double* src = P.kseg + (size_t)kr.y * P.ncols + col;
        double v = 0.0;
        for (int j0 = 0; j0 < kr.z; j0 += U) {
            double s[U];
#pragma unroll
            for (int u = 0; u < U; ++u) s[u] = j0 + u < kr.z ? src[(size_t)(j0 + u) * P.ncols] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j0 + u < kr.z) v += s[u];
        }
        const int r = sh_bf[kr.x];
        P.Kpart[(size_t)r * (n + 1) + P.collist[col]] = v;
    }
}

void hfk_launch_jk_part(const JKPart& P, int n, const int* sh_bf, const int* sh_bfp, const int* pair_A,
                        const int* pair_B, hipStream_t s)
{
    const long long total = (long long)P.nbra * P.na * P.nb + (P.same ? 0 : (long long)P.nket * P.nc * P.nd);
    const int width = 64 * ((P.ncols + 63) / 64);
    const int G = width >= 256 ? 1 : 256 / width;
    const long long blocks = (P.nkjob + G - 1) / G + (total + 255) / 256;
    if (blocks > 0)
        hipLaunchKernelGGL(k_jk_part, dim3((unsigned)blocks), dim3(256), 0, s, P, n, sh_bf, sh_bfp, pair_A, pair_B);
    const long long nt = (long long)P.nkrow * P.ncols;
    if (nt > 0) hipLaunchKernelGGL(k_krow_final, dim3(grid_for(nt, 256)), dim3(256), 0, s, P, n, sh_bf);
}

// G_ij = (J'_ij + J'_ji) - 1/2 (K'_ij + K'_ji), J' and K' the sums over the
// launches' partials in plan order.
__global__ void k_symmetrize_parts(int n, const double* Jpart, const double* Kpart, int nlaunch, const double* H,
                                   double* G_flat)
{
    const long long m = (long long)n * (n + 1) / 2;
    const long long np1 = (long long)(n + 1) * (n + 1);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        const double nn = (double)n + 0.5;
        long long i = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
        if (i < 0) i = 0;
        while (i > 0 && i * n - i * (i - 1) / 2 > idx) --i;
        while ((i + 1) * n - (i + 1) * i / 2 <= idx) ++i;
        const long long j = i + (idx - (i * n - i * (i - 1) / 2));
        const long long ld = n + 1;
        double jij = 0.0, jji = 0.0, kij = 0.0, kji = 0.0;
        for (int e = 0; e < nlaunch; ++e) {
            const double* Je = Jpart + (size_t)e * np1;
            const double* Ke = Kpart + (size_t)e * np1;
            jij += Je[i * ld + j];
            jji += Je[j * ld + i];
            kij += Ke[i * ld + j];
            kji += Ke[j * ld + i];
        }
        double g = (jij + jji) - 0.5 * (kij + kji);
        if (H) g += H[idx];
        G_flat[idx] = g;
    }
}

void hfk_launch_symmetrize_parts(int n, const double* Jpart, const double* Kpart, int nlaunch, const double* H,
                                 double* G_flat, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_symmetrize_parts, dim3(grid_for(m, 256)), dim3(256), 0, s, n, Jpart, Kpart, nlaunch, H, G_flat);
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

void hfk_launch_fock_prepare(int n, const double* D_flat, double* D, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2 + n + 1;
    hipLaunchKernelGGL(k_fock_prepare, dim3(grid_for(m, 256)), dim3(256), 0, s, n, D_flat, D);
}

void hfk_launch_full_to_flat(int n, const double* full, double* flat, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_full_to_flat, dim3(grid_for(m, 256)), dim3(256), 0, s, n, full, flat);
}
