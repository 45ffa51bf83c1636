// FP64 dense kernels for the SCF driver: an LDS-tiled GEMM on the CDNA
// FP64 matrix cores (v_mfma_f64_16x16x4f64) and small elementwise helpers.
#include <hip/hip_runtime.h>
#include <math.h>

#include "hfk_internal.hpp"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// 32 x 32 tiles: n = 832 gives 676 workgroups for 256 CUs (64 x 64 tiles gave
// 169; X^T F X at n = 832 took 0.364 ms with them, 0.088 ms now)
constexpr int TM = 32, TN = 32, TK = 32;
constexpr int WB = TM / 32;   // MFMA 16x16 blocks per wave along m and n
static_assert(TM == TN && TM % 32 == 0, "square tiles of a multiple of 32");

// C[m][n] = alpha * sum_k op(A)[m][k] op(B)[k][n] + beta * C[m][n], row-major.
// 256 threads = 4 waves; each wave owns a quarter (TM/2 x TN/2) of the tile
// as WB x WB MFMA 16x16 blocks.  MFMA f64 16x16x4 operand layout: lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; accumulator element r of lane l is
// D[l/16 + 4r][l%16].
template <int TA, int TB>
__global__ void __launch_bounds__(256) k_dgemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                               int lda, const double* __restrict__ B, int ldb, double beta,
                                               double* __restrict__ C, int ldc)
{
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
    v4d acc[WB][WB];
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    // the next k-tile's elements are loaded into registers while the MFMAs
    // work on the current one in LDS
    constexpr int EA = TM * TK / 256, EB = TN * TK / 256;
    static_assert(EA * 256 == TM * TK && EB * 256 == TN * TK, "whole tiles per workgroup");
    double ra[EA], rb[EB];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + 256 * q;
            int mm, kk;
            if (TA) { mm = e % TM; kk = e / TM; } else { kk = e % TK; mm = e / TK; }
            const int gm = m0 + mm, gk = k0 + kk;
            ra[q] = (gm < M && gk < K) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < EB; ++q) {
            const int e = tid + 256 * q;
            int nn, kk;
            if (TB) { kk = e % TK; nn = e / TK; } else { nn = e % TN; kk = e / TN; }
            const int gn = n0 + nn, gk = k0 + kk;
            rb[q] = (gn < N && gk < K) ? (TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn]) : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + 256 * q;
            if (TA) As[e / TM][e % TM] = ra[q];
            else As[e % TK][e / TK] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < EB; ++q) {
            const int e = tid + 256 * q;
            if (TB) Bs[e % TK][e / TK] = rb[q];
            else Bs[e / TN][e % TN] = rb[q];
        }
        __syncthreads();
        if (k0 + TK < K) load(k0 + TK);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            double a[WB], b[WB];
#pragma unroll
            for (int i = 0; i < WB; ++i) a[i] = As[kk + (lane >> 4)][wm + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < WB; ++j) b[j] = Bs[kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < WB; ++i)
#pragma unroll
                for (int j = 0; j < WB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + (lane >> 4) + 4 * r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    double* p = C + (size_t)gm * ldc + gn;
                    const double old = (beta != 0.0) ? beta * *p : 0.0;
                    *p = alpha * acc[i][j][r] + old;
                }
            }
}

// Scale columns (rows when pow_mode & 4) of U by f(w[k]):
// f = w^-1/2 (mode 0), f = w (mode 1).
__global__ void k_scale_cols(int n, double* U, const double* w, int pow_mode)
{
    const long long nn = (long long)n * n;
    const bool rows = (pow_mode & 4) != 0;
    const int mode = pow_mode & 3;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
        const int k = rows ? (int)(t / n) : (int)(t % n);
        const double f = (mode == 0) ? 1.0 / sqrt(w[k]) : w[k];
        U[t] *= f;
    }
}

// Two-stage deterministic reductions: per-block partials, then one block.
constexpr int RB = 256;

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__device__ __forceinline__ double block_max(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

// mode 0: sum x*y ; mode 1: max |x|
__global__ void __launch_bounds__(RB) k_reduce_partial(long long n, const double* x, const double* y, int mode, double* part)
{
    __shared__ double sh[RB];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
        v = (mode == 0) ? v + x[i] * y[i] : fmax(v, fabs(x[i]));
    v = (mode == 0) ? block_sum(v, sh) : block_max(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(RB) k_reduce_final(int nb, const double* part, int mode, double* out)
{
    __shared__ double sh[RB];
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) v = (mode == 0) ? v + part[i] : fmax(v, part[i]);
    v = (mode == 0) ? block_sum(v, sh) : block_max(v, sh);
    if (threadIdx.x == 0) *out = v;
}

// y = a*x + b*y
__global__ void k_axpby(long long n, double a, const double* x, double b, double* y)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

// E = T - T^T
__global__ void k_antisym(int n, const double* T, double* E)
{
    const long long nn = (long long)n * n;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t / n), j = (int)(t % n);
        E[t] = T[t] - T[(size_t)j * n + i];
    }
}

// Flat density with 2 on the first occ diagonal entries (Harris step input).
__global__ void k_guess_density(int n, int occ, double* Dflat)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += (long long)gridDim.x * blockDim.x)
        Dflat[t] = 0.0;
    __syncthreads();
}

__global__ void k_guess_diag(int n, int occ, double* Dflat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < occ && i < n) {
        const long long pos = (long long)i * n - (long long)i * (i - 1) / 2;   // (i,i)
        Dflat[pos] = 2.0;
    }
}

unsigned grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (unsigned)g;
}

}  // namespace

void hfk_launch_dgemm(int m, int n, int k, double alpha, const double* A, int lda, int ta, const double* B, int ldb,
                      int tb, double beta, double* C, int ldc, hipStream_t s)
{
    dim3 grid((n + TN - 1) / TN, (m + TM - 1) / TM);
    if (!ta && !tb) hipLaunchKernelGGL((k_dgemm<0, 0>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (ta && !tb) hipLaunchKernelGGL((k_dgemm<1, 0>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (!ta && tb) hipLaunchKernelGGL((k_dgemm<0, 1>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else hipLaunchKernelGGL((k_dgemm<1, 1>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void hfk_launch_scale_cols(int n, double* U, const double* w, int pow_mode, hipStream_t s)
{
    const long long nn = (long long)n * n;
    hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(nn, 256)), dim3(256), 0, s, n, U, w, pow_mode);
}

// out (device scalar) = sum x*y (mode 0) or max |x| (mode 1); part: >= 1024 doubles
void hfk_launch_reduce(long long n, const double* x, const double* y, int mode, double* part, double* out, hipStream_t s)
{
    unsigned g = grid_for(n, RB);
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_reduce_partial, dim3(g), dim3(RB), 0, s, n, x, y, mode, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RB), 0, s, (int)g, part, mode, out);
}

void hfk_launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, x, b, y);
}

void hfk_launch_antisym(int n, const double* T, double* E, hipStream_t s)
{
    hipLaunchKernelGGL(k_antisym, dim3(grid_for((long long)n * n, 256)), dim3(256), 0, s, n, T, E);
}

void hfk_launch_guess_density(int n, int occ, double* Dflat, hipStream_t s)
{
    const long long m = (long long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_guess_density, dim3(grid_for(m, 256)), dim3(256), 0, s, n, occ, Dflat);
    hipLaunchKernelGGL(k_guess_diag, dim3((occ + 255) / 256 + 1), dim3(256), 0, s, n, occ, Dflat);
}
