// Helpers shared by the quartet kernels (generated per-lane classes and the
// cooperative kernel for classes with d shells): quartet decoding, the
// pair-level selection (reference screen, Schwarz cut, precision split)
// and the canonical flat-array position of calcIntegrals.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

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

// Ket primitive-pair records as the generated ERI code reads them:
// prim(ik) is primitive pair ik of the ket, at(pk, f) its field f and
// q(pk, z) its screen value for Cartesian component z.
//  KetAoS: the pair's own records (PP_STRIDE doubles each) and qcomp[].
//  KetSoA: the position-major copy of the ket class (KetRun): field f of
//          primitive ik at base[(ik * NF + f) * len], the screen values
//          after the PP_STRIDE record fields, so that lanes holding
//          consecutive positions of one segment read consecutive words.
struct KetAoS {
    const double* p;
    const double* qcomp;
    __device__ __forceinline__ const double* prim(int ik) const { return p + (size_t)ik * PP_STRIDE; }
    __device__ __forceinline__ double at(const double* pk, int f) const { return pk[f]; }
    __device__ __forceinline__ double q(const double* pk, int z) const { return qcomp[(long long)pk[PP_QOFF] + z]; }
};
template <int NF>
struct KetSoA {
    const double* p;
    int len;
    __device__ __forceinline__ const double* prim(int ik) const { return p + (size_t)ik * NF * len; }
    __device__ __forceinline__ double at(const double* pk, int f) const { return pk[(size_t)f * len]; }
    __device__ __forceinline__ double q(const double* pk, int z) const { return pk[(size_t)(PP_STRIDE + z) * len]; }
};

// Pair-level selection of the per-lane kernels from the ket position's
// copies of pair_qhi and pair_sch (the same tests as quartet_selected).
__device__ __forceinline__ bool quartet_selected_pos(const QuartetArgs& a, double bqhi, double bsch, int pos)
{
    if (a.qtau > 0.0 && bqhi * a.kpos_q[pos] < a.qtau) return false;
    if (a.sch_cut > 0.0 || a.prec_sel != 0) {
        const double bound = bsch * a.kpos_q[a.nket + pos];
        if (bound < a.sch_cut) return false;
        if (a.prec_sel == 1 && !(bound >= a.fp32_thr)) return false;
        if (a.prec_sel == 2 && !(bound < a.fp32_thr)) return false;
    }
    return true;
}

// Kets of a task: its runs [x, y) of positions in the ket class list.
__device__ __forceinline__ int task_ket_count(const QuartetArgs& a, const int4& task)
{
    int n = 0;
    for (int r = task.y; r < task.z; ++r) n += a.runs[r].y - a.runs[r].x;
    return n;
}

// Position of the task's v-th ket (0 <= v < task_ket_count).
__device__ __forceinline__ int task_ket_pos(const QuartetArgs& a, const int4& task, int v)
{
    for (int r = task.y; r < task.z; ++r) {
        const KetRun q = a.runs[r];
        const int len = q.y - q.x;
        if (v < len) return q.x + v;
        v -= len;
    }
    return -1;
}

// End of the task's ket primitive range for ket pair kp (k1 = 255: all).
__device__ __forceinline__ int task_kprim_end(const QuartetArgs& a, int k1, int kp)
{
    return min(k1, a.pair_npp[kp]);
}

__device__ __forceinline__ bool quartet_selected(const QuartetArgs& a, int bp, int kp)
{
    if (a.qtau > 0.0 && a.pair_qhi[bp] * a.pair_qhi[kp] < a.qtau) return false;
    if (a.sch_cut > 0.0 || a.prec_sel != 0) {
        const double bound = a.pair_sch[bp] * a.pair_sch[kp];
        if (bound < a.sch_cut) return false;
        if (a.prec_sel == 1 && !(bound >= a.fp32_thr)) return false;
        if (a.prec_sel == 2 && !(bound < a.fp32_thr)) return false;
    }
    return true;
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
    const long long cnt = (long long)b0 - 1 - a0;   // sum_{k'=a0+1}^{b0-1} (n - k')
    r += cnt * n - ((long long)(b0 - 1) * b0 / 2 - (long long)a0 * (a0 + 1) / 2);
    return r + (b1 - b0);
}

constexpr int WAVE = 64;

}  // namespace
