// Device-side definitions shared by the hfk kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Primitive-pair record: one per (primitive of shell A, primitive of shell B)
// of a shell pair, 16 doubles (128 B, one cache line).
enum {
    PP_P = 0,       // p = a + b
    PP_PX, PP_PY, PP_PZ,   // Gaussian product centre P
    PP_AX, PP_AY, PP_AZ,   // P - A
    PP_BX, PP_BY, PP_BZ,   // P - B
    PP_K,           // c_a c_b exp(-ab/p |A-B|^2) / p   (base coefficients)
    PP_QLO,         // min over components of the reference screen value (NaN -> +inf)
    PP_QHI,         // max over components                             (NaN -> +inf)
    PP_QOFF,        // offset of this pair's per-component screen values in qcomp[]
    PP_OO2P,        // 1 / (2p)
    PP_EA,          // exponent of the first shell's primitive (gradient)
    PP_STRIDE = 16
};

#define HFK_TWO_PI_2_5 34.986836655249725   // 2 pi^(5/2)
#define HFK_PI 3.141592653589793

// Reference Boys grid (NumericalTools/TableBoys.hs:58-63): m = 0..14, x = 0.1 i, i = 0..1000.
// Stored node-major: the BOYS_REF_STRIDE values F_0..F_14 (+pad) of one node
// share one 128-byte line, so a Taylor step reads one or two lines.
#define BOYS_REF_M 15
#define BOYS_NX 1001
#define BOYS_REF_STRIDE 16
// Accurate grid: m = 0..ACC_M-1, x = 0.1 i, i = 0..ACC_NX-1 (x <= 50), node-major;
// above, asymptotic + recursion
#define BOYS_ACC_M 26
#define BOYS_ACC_NX 501
#define BOYS_ACC_STRIDE 32
#define BOYS_ACC_XMAX 50.0

struct BoysTables {
    const double* ref64;   // [BOYS_NX][BOYS_REF_STRIDE]
    const float* ref32;
    const double* acc64;   // [BOYS_ACC_NX][BOYS_ACC_STRIDE]
    const float* acc32;
    double asym[24];       // a_m = facOdd(floor(2m-1)) / 2^(m+1), with the reference's Int64 wrap
};

template <typename T> __device__ __forceinline__ const T* boys_ref_tab(const BoysTables& bt);
template <> __device__ __forceinline__ const double* boys_ref_tab<double>(const BoysTables& bt) { return bt.ref64; }
template <> __device__ __forceinline__ const float* boys_ref_tab<float>(const BoysTables& bt) { return bt.ref32; }
template <typename T> __device__ __forceinline__ const T* boys_acc_tab(const BoysTables& bt);
template <> __device__ __forceinline__ const double* boys_acc_tab<double>(const BoysTables& bt) { return bt.acc64; }
template <> __device__ __forceinline__ const float* boys_acc_tab<float>(const BoysTables& bt) { return bt.acc32; }

__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }

// F_m(x) for m = 0..L.
//  BOYS == 0: the reference's 6-term Taylor step from the 0.1-spaced node
//             table, asymptotic formula evaluated at the node above x_k = 100
//             (TableBoys.hs:45-51,65-69; Boys.hs:85-87).
//  BOYS == 1: accurate: 8-term Taylor about the nearest 0.1 node (|dx| <= 0.05)
//             of an accurate table for x < 50; F_0 = sqrt(pi/x)/2 and upward
//             recursion above.
template <typename T, int L, int BOYS>
__device__ __forceinline__ void boys_eval(T x, const BoysTables& bt, T* F)
{
    if constexpr (BOYS == 0) {
        const T fl = floor(x / T(0.1));
        const T xk = fl * T(0.1);
        const T dx = x - xk;
        T f[L + 6];
        if (xk > T(100)) {
            const T inv = T(1) / xk;
            T s = sqrt_t(T(HFK_PI) * inv);
#pragma unroll
            for (int m = 0; m < L + 6; ++m) {
                f[m] = T(bt.asym[m]) * s;
                s *= inv;
            }
        } else {
            int idx = (int)fl;
            idx = idx < 0 ? 0 : (idx > BOYS_NX - 1 ? BOYS_NX - 1 : idx);   // NaN/negative guard
            const T* tab = boys_ref_tab<T>(bt) + (size_t)idx * BOYS_REF_STRIDE;
#pragma unroll
            for (int m = 0; m < L + 6; ++m) f[m] = tab[m];
        }
        const T d2 = dx * dx, d3 = d2 * dx, d4 = d2 * d2, d5 = d4 * dx;
#pragma unroll
        for (int m = 0; m <= L; ++m)
            F[m] = f[m] - f[m + 1] * dx + (T(0.5) * f[m + 2]) * d2 - (T(1.0 / 6.0) * f[m + 3]) * d3
                   + (T(1.0 / 24.0) * f[m + 4]) * d4 - (T(1.0 / 120.0) * f[m + 5]) * d5;
    } else {
        if (x < T(BOYS_ACC_XMAX)) {
            const T fl = rint(x * T(10));
            const T dx = x - fl * T(0.1);
            int idx = (int)fl;
            idx = idx < 0 ? 0 : (idx > BOYS_ACC_NX - 1 ? BOYS_ACC_NX - 1 : idx);
            const T* tab = boys_acc_tab<T>(bt) + (size_t)idx * BOYS_ACC_STRIDE;
            const T mdx = -dx;
#pragma unroll
            for (int m = 0; m <= L; ++m) {
                T acc = tab[m + 7];
#pragma unroll
                for (int k = 7; k >= 1; --k) acc = tab[m + k - 1] + (mdx / T(k)) * acc;
                F[m] = acc;
            }
        } else {
            const T inv = T(1) / x;
            const T e = (T)exp((double)-x);
            F[0] = T(0.5) * sqrt_t(T(HFK_PI) * inv);
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = (T(2 * m + 1) * F[m] - e) * (T(0.5) * inv);
        }
    }
}

// Atomic add of a double with the hardware FP64 atomic (ds_add_f64 on an
// LDS address, the L2 atomic on a global one).
__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }
// The same, for LDS accumulators of one workgroup (the J/K kernels use no
// global FP atomics).
__device__ __forceinline__ void atomic_add_lds(double* p, double v) { unsafeAtomicAdd(p, v); }

// floor(x / 0.1) of the reference (TableBoys.hs:47-48) without a division:
// x*10 and x/0.1 differ by less than 2^-51 relative, so their floors can
// differ only when y = x*10 lies within y*2^-49 of an integer.  The cell
// matters only up to y = 1100 (x_k <= 100 decides between the table and the
// asymptote; above, x itself is used), where that distance is below 2e-12;
// arguments whose fractional part lies within 2e-12 of an integer take the
// exact division.  Beyond y = 1100 a cell off by one changes nothing.
__device__ __forceinline__ double boys_ref_cell(double x)
{
    const double y = x * 10.0;
    const double fl = floor(y);
    const double fr = y - fl;   // exact
    return (fabs(fr - 0.5) >= 0.5 - 2e-12) ? floor(x / 0.1) : fl;
}
// FP32: the same with x*10 and x/0.1 within 2^-22 relative (1100 * 2^-20 < 1.1e-3)
__device__ __forceinline__ float boys_ref_cell(float x)
{
    const float y = x * 10.0f;
    const float fl = floorf(y);
    const float fr = y - fl;
    return (fabsf(fr - 0.5f) >= 0.5f - 1.1e-3f) ? floorf(x / 0.1f) : fl;
}

// 1/sqrt(x): hardware estimate refined by Newton steps to full precision.
__device__ __forceinline__ double rsqrt_full(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - (0.5 * x) * (y * y));
    y = y * (1.5 - (0.5 * x) * (y * y));
    return y;
}
__device__ __forceinline__ float rsqrt_full(float x)
{
    float y = __builtin_amdgcn_rsqf(x);
    y = y * (1.5f - (0.5f * x) * (y * y));
    return y;
}

// F_0..F_L for the per-lane kernels, with as few instructions as possible.
//  BOYS == 0: x_k <= 100: the reference's 6-term Taylor step from the node
//             table (the same terms, summed with precomputed weights
//             w_k = (-dx)^k / k!); x_k > 100: the reference expands the
//             asymptotic form a_m sqrt(pi / x^(2m+1)) about x_k; here it is
//             evaluated at x itself (the 6-term expansion reproduces it to
//             ~1e-20 relative for x_k > 100, dx < 0.1).
//  BOYS == 1: as boys_eval.
template <typename T, int L, int BOYS>
__device__ __forceinline__ void boys_fast(T x, const BoysTables& bt, T* F)
{
    if constexpr (BOYS == 0) {
        const T fl = boys_ref_cell(x);
        const T xk = fl * T(0.1);
        if (xk <= T(100)) {
            int idx = (int)fl;
            idx = idx < 0 ? 0 : idx;   // NaN guard
            const T* tab = boys_ref_tab<T>(bt) + (unsigned)idx * BOYS_REF_STRIDE;
            T f[L + 6];
#pragma unroll
            for (int m = 0; m < L + 6; ++m) f[m] = tab[m];
            // sum_k (-dx)^k / k! f_{m+k}, k = 0..5, in Horner form with the
            // shared factors d/k
            const T d = xk - x;   // -dx
            const T d2 = T(0.5) * d, d3 = T(1.0 / 3.0) * d, d4 = T(0.25) * d, d5 = T(0.2) * d;
#pragma unroll
            for (int m = 0; m <= L; ++m)
                F[m] = f[m] + d * (f[m + 1] + d2 * (f[m + 2] + d3 * (f[m + 3] + d4 * (f[m + 4] + d5 * f[m + 5]))));
        } else {
            const T rs = rsqrt_full(x * T(1.0 / HFK_PI));   // sqrt(pi / x)
            const T inv = (rs * rs) * T(1.0 / HFK_PI);      // 1 / x
            T s = rs;
#pragma unroll
            for (int m = 0; m <= L; ++m) {
                F[m] = T(bt.asym[m]) * s;
                s *= inv;
            }
        }
    } else {
        boys_eval<T, L, BOYS>(x, bt, F);
    }
}


// Geometry of one primitive quartet from the bra record pb and the ket
// pair (q, Q): alpha = pq/(p+q), R_PQ, T = alpha |R_PQ|^2, rs = 1/sqrt(p+q).
template <typename T>
__device__ __forceinline__ void quartet_geom(const double* pb, T kq, T kQx, T kQy, T kQz, T& alpha, T& X, T& Y, T& Z,
                                             T& rs, T& Targ)
{
    const T bp = T(pb[PP_P]);
    const T pq = bp + kq;
    rs = rsqrt_full(pq);
    alpha = (bp * kq) * (rs * rs);
    X = T(pb[PP_PX]) - kQx;
    Y = T(pb[PP_PY]) - kQy;
    Z = T(pb[PP_PZ]) - kQz;
    Targ = alpha * ((X * X + Y * Y) + Z * Z);
}
