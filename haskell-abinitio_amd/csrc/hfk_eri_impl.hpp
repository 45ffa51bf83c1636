// Shell-quartet kernels: ERI evaluation (generated McMurchie-Davidson code)
// fused with J/K digestion, flat-array export or the Schwarz diagonal.
// Included by one generated translation unit per angular-momentum class,
// which first defines eri_class<LA,LB,LC,LD,T,BOYS,MODE> and afterwards
// instantiates dispatch<LA,LB,LC,LD>.
//
// Screening follows the reference exactly (IntegralsEvaluation.hs:305-315):
// a primitive quartet contributes to a Cartesian component (ab|cd) iff
// !(q_ab q_cd < tau).  MODE 1 applies that screen (branch-free selects on
// per-component values loaded once per primitive pair); MODE 0 evaluates
// unscreened (screen off, or the Schwarz diagonal).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "hfk_quartet_common.hpp"

namespace {

// Primitive-pair sub-ranges [b0,b1) of the bra and [k0,k1) of the ket
// (b1/k1 < 0: to the end); bra records read from global memory.
template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
__device__ __forceinline__ void eval_quartet(const QuartetArgs& a, int bp, int kp, T* V, int b0 = 0, int b1 = -1,
                                             int k0 = 0, int k1 = -1)
{
    const double* bra = a.pp + (size_t)(a.pair_pp0[bp] + b0) * PP_STRIDE;
    const double* ket = a.pp + (size_t)(a.pair_pp0[kp] + k0) * PP_STRIDE;
    const int nbp = (b1 < 0 ? a.pair_npp[bp] : b1) - b0;
    const int nkp = (k1 < 0 ? a.pair_npp[kp] : k1) - k0;
    const double* bq = a.qcomp + (MODE == 1 ? (long long)bra[PP_QOFF] : 0LL);
    eri_class<LA, LB, LC, LD, T, BOYS, MODE>(bra, bq, nbp, ket, nkp, a.qcomp, a.bt, V, a.qtau);
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void write_flat(const QuartetArgs& a, int bp, int kp, const double* V)
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
                    a.flat[flat_pos(a, i, j, k, l)] = V[idx] * (si * sj * sk * sl);
                }
            }
        }
    }
}

// Thread-per-quartet kernel for the flat export and the Schwarz diagonal.
template <int LA, int LB, int LC, int LD, int BOYS, int OP>
__global__ void __launch_bounds__(128) k_quartets(QuartetArgs a)
{
    constexpr int NV = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.nquart; q += stride) {
        int bp, kp;
        if (OP == HFK_OP_DIAG) {
            bp = kp = a.bra_list[q];
        } else {
            long long i, j;
            decode_quartet(a, q, i, j);
            bp = a.bra_list[i];
            kp = a.ket_list[j];
            if (!quartet_selected(a, bp, kp)) continue;
        }
        double V[NV];
        if (a.qtau > 0.0) eval_quartet<LA, LB, LC, LD, double, BOYS, 1>(a, bp, kp, V);
        else eval_quartet<LA, LB, LC, LD, double, BOYS, 0>(a, bp, kp, V);
        if constexpr (OP == HFK_OP_FLAT) {
            write_flat<LA, LB, LC, LD>(a, bp, kp, V);
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
                    m = fmax(m, fabs(V[k * NBRA + k]) * s * s);
                }
            a.diag[bp] = m;
        }
    }
}

template <int LA, int LB, int LC, int LD, int BOYS, int OP>
void launch_quartets_one(const QuartetArgs& a, hipStream_t s)
{
    if (a.nquart <= 0) return;
    const int block = 128;
    long long nb = (a.nquart + block - 1) / block;
    const long long cap = 256LL * 8 * 16;
    if (nb > cap) nb = cap;
    hipLaunchKernelGGL((k_quartets<LA, LB, LC, LD, BOYS, OP>), dim3((unsigned)nb), dim3(block), 0, s, a);
}

// ---------------------------------------------------------------------
// Tiled J/K kernel: one workgroup per task = (bra shell pair, runs of the ket
// class list (each of one primitive-pair count), primitive-pair ranges of
// bra and ket).  Each lane loops over its ket primitive pairs
// (vector loads, once each) and, inside, over the bra primitive pairs
// (wave-uniform, scalar loads).  Splitting heavy quartets into primitive
// ranges evens out the work per lane; the digestion is linear in the
// integrals, so partial quartets are digested independently.  The bra pair is
// wave-uniform (its data comes through scalar loads); each lane takes ket
// pairs of the range.  D, J', K' are (n+1) x (n+1) with a zero row/column
// n that absent Cartesian components point to, so the digestion has no
// branches.  Per quartet:
//   J'_ab += 2f sum_cd V D_cd      lane registers, one wave reduction per task
//   K'_a. , K'_b.                  rows of the two bra shells staged in LDS
//                                  (ds_add_f64), flushed once per task
//   J'_cd += 2f sum_ab V D_ab      one global FP64 atomic per ket function pair,
//                                  by ket position (coalesced), scattered into J' later
// ---------------------------------------------------------------------
// Workgroups are 64 lanes, or 256 lanes (4 waves sharing one set of staged K
// rows) when the rows are long enough to limit occupancy through LDS.
template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
__global__ void __launch_bounds__(256) k_jk_tiles(QuartetArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NKET = NC * ND, NBRA = NA * NB;
    extern __shared__ double krow[];   // [(NA+NB)][ld] when lds_rows, then the staged bra
    const int4 task = a.tasks[blockIdx.x];
    const int bp = task.x;
    // primitive-pair ranges: bra [b0,b1), ket [k0,k1), 8 bits each
    const int bprim0 = (task.w >> 24) & 0xff, bprim1 = (task.w >> 16) & 0xff;
    const int kprim0 = (task.w >> 8) & 0xff, kprim1 = task.w & 0xff;
    const int ld = a.n + 1;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & (WAVE - 1);
    const bool lds = a.lds_rows != 0;
    if (lds)
        for (int t = tid; t < (NA + NB) * ld; t += nthr) krow[t] = 0.0;
    // the task's bra primitive-pair records and their per-component screen
    // values, read by every lane in every ket iteration: staged once in LDS
    const int nbs = bprim1 - bprim0;
    double* sbra = krow + (lds ? (NA + NB) * ld : 0);
    double* sbq = sbra + nbs * PP_STRIDE;
    {
        const double* g = a.pp + (size_t)(a.pair_pp0[bp] + bprim0) * PP_STRIDE;
        for (int t = tid; t < nbs * PP_STRIDE; t += nthr) sbra[t] = g[t];
        if constexpr (MODE == 1) {
            const double* q = a.qcomp + (long long)g[PP_QOFF];
            for (int t = tid; t < nbs * NBRA; t += nthr) sbq[t] = q[t];
        }
    }
    __syncthreads();

    const int A = a.pair_A[bp], B = a.pair_B[bp];
    int ia[NA], ib[NB];
    double sa[NA], sb[NB];
#pragma unroll
    for (int x = 0; x < NA; ++x) { ia[x] = a.sh_bfp[A * HFK_MAXCOMP + x]; sa[x] = a.sh_scale[A * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < NB; ++x) { ib[x] = a.sh_bfp[B * HFK_MAXCOMP + x]; sb[x] = a.sh_scale[B * HFK_MAXCOMP + x]; }
    const double* __restrict__ D = a.D;
    double Dab[NBRA];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) Dab[x * NB + y] = D[(size_t)ia[x] * ld + ib[y]];
    double jab[NBRA];
#pragma unroll
    for (int k = 0; k < NBRA; ++k) jab[k] = 0.0;

    // ket-side J'_cd: summed by position in the ket class list, so the
    // lanes' atomics fall on consecutive addresses; tasks resident at the same
    // time share ket runs, so workgroups spread them over a few copies
    double* __restrict__ Jk = a.Jk + (size_t)(blockIdx.x % (unsigned)a.jk_copies) * a.nket * NKET;
    // (lanes take the kets in run order: a wave meets one primitive-pair
    // count except at run boundaries)
    const int len = task_ket_count(a, task);
    for (int it = tid; it < len; it += nthr) {
        const int pos = task_ket_pos(a, task, it);
        const int kp = a.ket_idx[pos];
        if (!quartet_selected(a, bp, kp)) continue;
        T Vt[NBRA * NKET];
        eri_class<LA, LB, LC, LD, T, BOYS, MODE>(sbra, sbq, nbs, a.pp + (size_t)(a.pair_pp0[kp] + kprim0) * PP_STRIDE,
                                                 task_kprim_end(a, kprim1, kp) - kprim0, a.qcomp, a.bt, Vt, a.qtau);
        const int C = a.pair_A[kp], Dd = a.pair_B[kp];
        int ic[NC], id[ND];
        double sc[NC], sd[ND];
#pragma unroll
        for (int x = 0; x < NC; ++x) { ic[x] = a.sh_bfp[C * HFK_MAXCOMP + x]; sc[x] = a.sh_scale[C * HFK_MAXCOMP + x]; }
#pragma unroll
        for (int x = 0; x < ND; ++x) { id[x] = a.sh_bfp[Dd * HFK_MAXCOMP + x]; sd[x] = a.sh_scale[Dd * HFK_MAXCOMP + x]; }
        double f = 1.0;
        if (A == B) f *= 0.5;
        if (C == Dd) f *= 0.5;
        if (bp == kp) f *= 0.5;
        double V[NBRA * NKET];
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int z = 0; z < NC; ++z)
#pragma unroll
                    for (int w = 0; w < ND; ++w) {
                        const int k = ((x * NB + y) * NC + z) * ND + w;
                        V[k] = (double)Vt[k] * (f * sa[x] * sb[y] * sc[z] * sd[w]);
                    }
        double Dcd[NKET];
#pragma unroll
        for (int z = 0; z < NC; ++z)
#pragma unroll
            for (int w = 0; w < ND; ++w) Dcd[z * ND + w] = D[(size_t)ic[z] * ld + id[w]];
#pragma unroll
        for (int b = 0; b < NBRA; ++b) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NKET; ++k) s += V[b * NKET + k] * Dcd[k];
            jab[b] += s;
        }
#pragma unroll
        for (int k = 0; k < NKET; ++k) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NBRA; ++b) s += V[b * NKET + k] * Dab[b];
            atomic_add_f64(&Jk[(size_t)pos * NKET + k], 2.0 * s);
        }
        // density blocks D_bd, D_bc, D_ad, D_ac
        double Dbd[NB * ND], Dbc[NB * NC], Dad[NA * ND], Dac[NA * NC];
#pragma unroll
        for (int y = 0; y < NB; ++y) {
#pragma unroll
            for (int w = 0; w < ND; ++w) Dbd[y * ND + w] = D[(size_t)ib[y] * ld + id[w]];
#pragma unroll
            for (int z = 0; z < NC; ++z) Dbc[y * NC + z] = D[(size_t)ib[y] * ld + ic[z]];
        }
#pragma unroll
        for (int x = 0; x < NA; ++x) {
#pragma unroll
            for (int w = 0; w < ND; ++w) Dad[x * ND + w] = D[(size_t)ia[x] * ld + id[w]];
#pragma unroll
            for (int z = 0; z < NC; ++z) Dac[x * NC + z] = D[(size_t)ia[x] * ld + ic[z]];
        }
        double* __restrict__ Kg = a.K;
#pragma unroll
        for (int x = 0; x < NA; ++x) {
#pragma unroll
            for (int z = 0; z < NC; ++z) {
                double s = 0.0;
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int w = 0; w < ND; ++w) s += V[((x * NB + y) * NC + z) * ND + w] * Dbd[y * ND + w];
                if (lds) atomic_add_f64(&krow[x * ld + ic[z]], s);
                else atomic_add_f64(&Kg[(size_t)ia[x] * ld + ic[z]], s);
            }
#pragma unroll
            for (int w = 0; w < ND; ++w) {
                double s = 0.0;
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * Dbc[y * NC + z];
                if (lds) atomic_add_f64(&krow[x * ld + id[w]], s);
                else atomic_add_f64(&Kg[(size_t)ia[x] * ld + id[w]], s);
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
                    for (int w = 0; w < ND; ++w) s += V[((x * NB + y) * NC + z) * ND + w] * Dad[x * ND + w];
                if (lds) atomic_add_f64(&krow[(NA + y) * ld + ic[z]], s);
                else atomic_add_f64(&Kg[(size_t)ib[y] * ld + ic[z]], s);
            }
#pragma unroll
            for (int w = 0; w < ND; ++w) {
                double s = 0.0;
#pragma unroll
                for (int x = 0; x < NA; ++x)
#pragma unroll
                    for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * Dac[x * NC + z];
                if (lds) atomic_add_f64(&krow[(NA + y) * ld + id[w]], s);
                else atomic_add_f64(&Kg[(size_t)ib[y] * ld + id[w]], s);
            }
        }
    }
    // J'_ab: wave reduction, one atomic per bra function pair per wave
#pragma unroll
    for (int b = 0; b < NBRA; ++b) {
        double v = jab[b];
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
        if (lane == 0 && v != 0.0) atomic_add_f64(&a.J[(size_t)ia[b / NB] * ld + ib[b % NB]], 2.0 * v);
    }
    // flush the staged K rows (row by row: no index division)
    if (lds) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NA + NB; ++r) {
            double* __restrict__ Kr = a.K + (size_t)(r < NA ? ia[r] : ib[r - NA]) * ld;
            const double* sr = krow + r * ld;
            for (int col = tid; col < ld; col += nthr) {
                const double v = sr[col];
                if (v != 0.0) atomic_add_f64(&Kr[col], v);
            }
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
void launch_tiles(const QuartetArgs& a, hipStream_t s)
{
    if (a.ntask <= 0) return;
    constexpr int NR = ncart(LA) + ncart(LB);
    const size_t shm = (a.lds_rows ? sizeof(double) * NR * (size_t)(a.n + 1) : 0) + hfk_bra_stage_bytes(LA, LB, a.bra_stage);
    long long done = 0;
    while (done < a.ntask) {   // grid.x limit
        const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
        QuartetArgs b = a;
        b.tasks = a.tasks + done;
        hipLaunchKernelGGL((k_jk_tiles<LA, LB, LC, LD, T, BOYS, MODE>), dim3((unsigned)cnt), dim3(a.block), shm, s, b);
        done += cnt;
    }
}

// op: HFK_OP_JK (mode 0 main, 1 correction), HFK_OP_FLAT, HFK_OP_DIAG.
template <int LA, int LB, int LC, int LD>
int dispatch(int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)
{
    if (op == HFK_OP_JK) {
#define HFK_T(T)                                                                        \
        if (boys == 0) {                                                                  \
            if (mode == 0) launch_tiles<LA, LB, LC, LD, T, 0, 0>(a, s);                   \
            else launch_tiles<LA, LB, LC, LD, T, 0, 1>(a, s);                             \
        } else {                                                                          \
            if (mode == 0) launch_tiles<LA, LB, LC, LD, T, 1, 0>(a, s);                   \
            else launch_tiles<LA, LB, LC, LD, T, 1, 1>(a, s);                             \
        }
        if (prec == 0) { HFK_T(double) } else { HFK_T(float) }
#undef HFK_T
        return 0;
    }
    if (prec != 0) return -1;
    if (op == HFK_OP_FLAT) {
        if (boys == 0) launch_quartets_one<LA, LB, LC, LD, 0, HFK_OP_FLAT>(a, s);
        else launch_quartets_one<LA, LB, LC, LD, 1, HFK_OP_FLAT>(a, s);
        return 0;
    }
    if (op == HFK_OP_DIAG) {
        if (boys == 0) launch_quartets_one<LA, LB, LC, LD, 0, HFK_OP_DIAG>(a, s);
        else launch_quartets_one<LA, LB, LC, LD, 1, HFK_OP_DIAG>(a, s);
        return 0;
    }
    return -1;
}

}  // namespace
