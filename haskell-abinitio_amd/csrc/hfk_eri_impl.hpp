// Shell-quartet kernels: ERI evaluation (generated McMurchie-Davidson code)
// fused with J/K digestion, flat-array export or Schwarz diagonal.  Included
// by one generated translation unit per angular-momentum class, which
// defines HFK_CLASS_HEADER (the generated eri_<class> function) and
// instantiates hfk_dispatch_<class>.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "hfk_internal.hpp"

namespace {

constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Decode linear quartet index -> (bra list position i, ket list position j).
__device__ __forceinline__ void decode_quartet(const QuartetArgs& a, long long q, long long& i, long long& j)
{
    if (a.same) {
        // q = i(i+1)/2 + j, 0 <= j <= i
        double x = sqrt(8.0 * (double)q + 1.0);
        long long ii = (long long)((x - 1.0) * 0.5);
        while (ii * (ii + 1) / 2 > q) --ii;
        while ((ii + 1) * (ii + 2) / 2 <= q) ++ii;
        i = ii;
        j = q - ii * (ii + 1) / 2;
    } else {
        i = q / a.nket;
        j = q - i * a.nket;
    }
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS>
__device__ __forceinline__ void eval_quartet(const QuartetArgs& a, int bp, int kp, T* V)
{
    const double* bra = a.pp + (size_t)a.pair_pp0[bp] * PP_STRIDE;
    const double* ket = a.pp + (size_t)a.pair_pp0[kp] * PP_STRIDE;
    const int nbp = a.pair_npp[bp], nkp = a.pair_npp[kp];
    eri_class<LA, LB, LC, LD, T, BOYS>(bra, nbp, ket, nkp, a.qcomp, a.bt, V, a.qtau);
}

template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void digest_jk(const QuartetArgs& a, int bp, int kp, const T* Vt)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NKET = NC * ND;
    const int A = a.pair_A[bp], B = a.pair_B[bp], C = a.pair_A[kp], Dd = a.pair_B[kp];
    int ia[NA], ib[NB], ic[NC], id[ND];
    double sa[NA], sb[NB], sc[NC], sd[ND];
#pragma unroll
    for (int x = 0; x < NA; ++x) { ia[x] = a.sh_bf[A * HFK_MAXCOMP + x]; sa[x] = a.sh_scale[A * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < NB; ++x) { ib[x] = a.sh_bf[B * HFK_MAXCOMP + x]; sb[x] = a.sh_scale[B * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < NC; ++x) { ic[x] = a.sh_bf[C * HFK_MAXCOMP + x]; sc[x] = a.sh_scale[C * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < ND; ++x) { id[x] = a.sh_bf[Dd * HFK_MAXCOMP + x]; sd[x] = a.sh_scale[Dd * HFK_MAXCOMP + x]; }

    double f = 1.0;
    if (A == B) f *= 0.5;
    if (C == Dd) f *= 0.5;
    if (bp == kp) f *= 0.5;
    const int n = a.n;
    const double* D = a.D;
    auto dget = [&](int p, int q) -> double { return (p < 0 || q < 0) ? 0.0 : D[(size_t)p * n + q]; };

    double V[NA * NB * NC * ND];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y)
#pragma unroll
            for (int z = 0; z < NC; ++z)
#pragma unroll
                for (int w = 0; w < ND; ++w) {
                    const int k = ((x * NB + y) * NC + z) * ND + w;
                    V[k] = (double)Vt[k] * (sa[x] * sb[y] * sc[z] * sd[w]);
                }

    // J'_ab += 2f sum_cd V D_cd ; J'_cd += 2f sum_ab V D_ab
    double Dcd[NC * ND], Dab[NA * NB];
#pragma unroll
    for (int z = 0; z < NC; ++z)
#pragma unroll
        for (int w = 0; w < ND; ++w) Dcd[z * ND + w] = dget(ic[z], id[w]);
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) Dab[x * NB + y] = dget(ia[x], ib[y]);
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NKET; ++k) s += V[(x * NB + y) * NKET + k] * Dcd[k];
            if (ia[x] >= 0 && ib[y] >= 0) atomic_add_f64(&a.J[(size_t)ia[x] * n + ib[y]], 2.0 * f * s);
        }
#pragma unroll
    for (int k = 0; k < NKET; ++k) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NA * NB; ++b) s += V[b * NKET + k] * Dab[b];
        const int z = k / ND, w = k % ND;
        if (ic[z] >= 0 && id[w] >= 0) atomic_add_f64(&a.J[(size_t)ic[z] * n + id[w]], 2.0 * f * s);
    }
    // K'_ac += f sum_bd V D_bd ; K'_ad += f sum_bc V D_bc
    // K'_bc += f sum_ad V D_ad ; K'_bd += f sum_ac V D_ac
#pragma unroll
    for (int x = 0; x < NA; ++x) {
#pragma unroll
        for (int z = 0; z < NC; ++z) {
            double s = 0.0;
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int w = 0; w < ND; ++w) s += V[((x * NB + y) * NC + z) * ND + w] * dget(ib[y], id[w]);
            if (ia[x] >= 0 && ic[z] >= 0) atomic_add_f64(&a.K[(size_t)ia[x] * n + ic[z]], f * s);
        }
#pragma unroll
        for (int w = 0; w < ND; ++w) {
            double s = 0.0;
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * dget(ib[y], ic[z]);
            if (ia[x] >= 0 && id[w] >= 0) atomic_add_f64(&a.K[(size_t)ia[x] * n + id[w]], f * s);
        }
    }
#pragma unroll
    for (int y = 0; y < NB; ++y) {
#pragma unroll
        for (int z = 0; z < NC; ++z) {
            double s = 0.0;
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int w = 0; w < ND; ++w) s += V[((x * NB + y) * NC + z) * ND + w] * dget(ia[x], id[w]);
            if (ib[y] >= 0 && ic[z] >= 0) atomic_add_f64(&a.K[(size_t)ib[y] * n + ic[z]], f * s);
        }
#pragma unroll
        for (int w = 0; w < ND; ++w) {
            double s = 0.0;
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * dget(ia[x], ic[z]);
            if (ib[y] >= 0 && id[w] >= 0) atomic_add_f64(&a.K[(size_t)ib[y] * n + id[w]], f * s);
        }
    }
}

// Canonical position of a function quartet in calcIntegrals' flat array
// (HartreeFock.hs:153-181): sortKeys, then the enumeration rank.
__device__ __forceinline__ long long flat_pos(const QuartetArgs& a, int i, int j, int k, int l)
{
    int a0 = min(i, j), a1 = max(i, j), b0 = min(k, l), b1 = max(k, l);
    if (!(a0 < b0 || (a0 == b0 && a1 <= b1))) {
        int t0 = a0, t1 = a1;
        a0 = b0; a1 = b1; b0 = t0; b1 = t1;
    }
    const long long n = a.n;
    long long r = a.flat_base[(size_t)a0 * n + a1];
    if (b0 == a0) return r + (b1 - a1);
    r += n - a1;
    // sum_{k'=a0+1}^{b0-1} (n - k')
    const long long cnt = (long long)b0 - 1 - a0;
    r += cnt * n - ((long long)(b0 - 1) * b0 / 2 - (long long)a0 * (a0 + 1) / 2);
    return r + (b1 - b0);
}

template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void write_flat(const QuartetArgs& a, int bp, int kp, const T* Vt)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    const int A = a.pair_A[bp], B = a.pair_B[bp], C = a.pair_A[kp], Dd = a.pair_B[kp];
#pragma unroll
    for (int x = 0; x < NA; ++x) {
        const int i = a.sh_bf[A * HFK_MAXCOMP + x];
        const double si = a.sh_scale[A * HFK_MAXCOMP + x];
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            const int j = a.sh_bf[B * HFK_MAXCOMP + y];
            const double sj = a.sh_scale[B * HFK_MAXCOMP + y];
#pragma unroll
            for (int z = 0; z < NC; ++z) {
                const int k = a.sh_bf[C * HFK_MAXCOMP + z];
                const double sk = a.sh_scale[C * HFK_MAXCOMP + z];
#pragma unroll
                for (int w = 0; w < ND; ++w) {
                    const int l = a.sh_bf[Dd * HFK_MAXCOMP + w];
                    const double sl = a.sh_scale[Dd * HFK_MAXCOMP + w];
                    if (i < 0 || j < 0 || k < 0 || l < 0) continue;
                    const int idx = ((x * NB + y) * NC + z) * ND + w;
                    a.flat[flat_pos(a, i, j, k, l)] = (double)Vt[idx] * (si * sj * sk * sl);
                }
            }
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int OP>
__global__ void __launch_bounds__(128) k_quartets(QuartetArgs a)
{
    constexpr int NV = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.nquart; q += stride) {
        long long i, j;
        int bp, kp;
        if (OP == HFK_OP_DIAG) {
            bp = kp = a.bra_list[q];
        } else {
            decode_quartet(a, q, i, j);
            bp = a.bra_list[i];
            kp = a.ket_list[j];
            if (a.qtau > 0.0 && a.pair_qhi[bp] * a.pair_qhi[kp] < a.qtau) continue;
            if (a.sch_cut > 0.0 || a.prec_sel != 0) {
                const double bound = a.pair_sch[bp] * a.pair_sch[kp];
                if (bound < a.sch_cut) continue;
                if (a.prec_sel == 1 && !(bound >= a.fp32_thr)) continue;
                if (a.prec_sel == 2 && !(bound < a.fp32_thr)) continue;
            }
        }
        T V[NV];
        eval_quartet<LA, LB, LC, LD, T, BOYS>(a, bp, kp, V);
        if constexpr (OP == HFK_OP_JK) {
            digest_jk<LA, LB, LC, LD, T>(a, bp, kp, V);
        } else if constexpr (OP == HFK_OP_FLAT) {
            write_flat<LA, LB, LC, LD, T>(a, bp, kp, V);
        } else {
            // Schwarz diagonal: max_ab |(ab|ab)| with component scales
            constexpr int NA = ncart(LA), NB = ncart(LB), NBRA = NA * NB;
            const int A = a.pair_A[bp], B = a.pair_B[bp];
            double m = 0.0;
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) {
                    const double s = a.sh_scale[A * HFK_MAXCOMP + x] * a.sh_scale[B * HFK_MAXCOMP + y];
                    const int k = x * NB + y;
                    m = fmax(m, fabs((double)V[k * NBRA + k]) * s * s);
                }
            a.diag[bp] = m;
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int OP>
void launch_one(const QuartetArgs& a, hipStream_t s)
{
    if (a.nquart <= 0) return;
    const int block = 128;
    long long nb = (a.nquart + block - 1) / block;
    const long long cap = 256LL * 8 * 16;   // enough waves to fill 256 CUs several times
    if (nb > cap) nb = cap;
    hipLaunchKernelGGL((k_quartets<LA, LB, LC, LD, T, BOYS, OP>), dim3((unsigned)nb), dim3(block), 0, s, a);
}

template <int LA, int LB, int LC, int LD>
int dispatch(int op, int prec, int boys, const QuartetArgs& a, hipStream_t s)
{
#define HFK_OPS(T, B)                                                                          \
    switch (op) {                                                                            \
    case HFK_OP_JK: launch_one<LA, LB, LC, LD, T, B, HFK_OP_JK>(a, s); break;                \
    case HFK_OP_FLAT: launch_one<LA, LB, LC, LD, T, B, HFK_OP_FLAT>(a, s); break;            \
    case HFK_OP_DIAG: launch_one<LA, LB, LC, LD, T, B, HFK_OP_DIAG>(a, s); break;            \
    default: return -1;                                                                      \
    }
    if (prec == 0) {
        if (boys == 0) { HFK_OPS(double, 0) } else { HFK_OPS(double, 1) }
    } else {
        if (op != HFK_OP_JK) return -1;
        if (boys == 0) { launch_one<LA, LB, LC, LD, float, 0, HFK_OP_JK>(a, s); }
        else { launch_one<LA, LB, LC, LD, float, 1, HFK_OP_JK>(a, s); }
    }
#undef HFK_OPS
    return 0;
}

}  // namespace
