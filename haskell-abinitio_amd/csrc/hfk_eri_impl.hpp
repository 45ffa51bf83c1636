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
    const KetAoS kr{ket, a.qcomp};
    eri_class<LA, LB, LC, LD, T, BOYS, MODE, 0>(bra, bq, nbp, kr, nkp, a.bt, V, a.qtau);
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
// bra and ket).  Each lane loops over its ket primitive pairs and, inside,
// over the bra primitive pairs (staged in LDS, wave-uniform).  Splitting
// heavy quartets into primitive ranges evens out the work per lane; the
// digestion is linear in the integrals, so partial quartets are digested
// independently.  D is (n+1) x (n+1) with a zero row/column n that absent
// Cartesian components point to, so the digestion has no branches.
//
// Lanes take consecutive positions of a run, and everything a lane reads per
// ket is laid out by ket position: the ket's primitive-pair records
// (position-major copy of the class), its descriptor (function indices,
// compact columns, flags) and screen values, so those loads are coalesced.
// The density elements D_ac, D_ad, D_bc, D_bd come from the bra shells' rows
// staged in LDS over the compact columns (drows); only D_cd is gathered.
// Per quartet:
//   J'_ab += 2f sum_cd V D_cd   lane registers; at the end one wave
//                               reduction, the waves summed in order, one
//                               store to the task's slot
//   J'_cd  = 2f sum_ab V D_ab   one store per ket to its own slot entry
//   K'_a. , K'_b.               rows of the two bra shells staged in LDS
//                               (compact columns: the functions of the
//                               launch's ket shells), stored to the task's
//                               slot at the end
// No global atomics: the slots are summed in a fixed order afterwards, so a
// Fock build is bitwise reproducible.  In a workgroup of several waves the
// waves add into the LDS rows in turn (a barrier between turns); within one
// wave the hardware orders the lanes' adds, so the LDS sums do not depend on
// scheduling either.
// ---------------------------------------------------------------------
// Waves per SIMD the register allocation must leave room for.  FP64: the
// L = 3 classes need a few registers over 256 (VGPRs + AGPRs), (ps|ps) one
// over 128, which alone would leave one and three waves per SIMD (a few bytes
// of spill per ket instead).  FP32 s/p classes (the FP32 part of a
// mixed-precision task list, launched without the density rows in LDS so
// that LDS does not cap the workgroups per CU): more resident waves pay for
// some spill, (ps|ss) 6, (ps|ps) 5, (pp|ps) 3, (pp|pp) 2; (pp|ss) spills too
// much at 5 and keeps the 4 its registers allow.
template <int LA, int LB, int LC, int LD, typename T>
constexpr int jk_min_waves()
{
    constexpr int L = LA + LB + LC + LD;
    constexpr bool pss = LA == 1 && LB == 0 && LC == 1 && LD == 0;
    if constexpr (sizeof(T) == 4 && LA <= 1)
        return L == 1 ? 6 : pss ? 5 : L == 3 ? 3 : L == 4 ? 2 : 1;
    return L == 3 ? 2 : pss ? 4 : 1;
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(jk_min_waves<LA, LB, LC, LD, T>())))
k_jk_tiles(QuartetArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NKET = NC * ND, NBRA = NA * NB;
    constexpr int NR = NA + NB;
    constexpr int NF = hfk_ket_fields(LC, LD);
    extern __shared__ double smem[];
    const long long tix = a.task0 + blockIdx.x;    // slot of this task
    const int4 task = a.tasks[blockIdx.x];
    const int bp = task.x;
    // primitive-pair ranges: bra [b0,b1), ket [k0,k1), 8 bits each
    const int bprim0 = (task.w >> 24) & 0xff, bprim1 = (task.w >> 16) & 0xff;
    const int kprim0 = (task.w >> 8) & 0xff, kprim1 = task.w & 0xff;
    const int ld = a.n + 1;
    const int ncols = a.ncols, ldk = ncols + 1;    // LDS row: compact columns + one discarded
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int nwaves = nthr / WAVE;
    const bool drows = a.drows != 0;
    double* krow = smem;                           // [NR][ldk] K' rows
    double* drow = krow + NR * ldk;                // [NR][ldk] D rows of the bra shells (drows)
    double* sred = drow + (drows ? NR * ldk : 0);  // [4][NBRA] cross-wave J'_ab
    long long* rjcd = (long long*)(sred + 4 * NBRA);   // [HFK_TASK_RUNS + 1] the task's runs: J'_cd entry,
    long long* rsoa = rjcd + HFK_TASK_RUNS + 1;        //   record base,
    int* rx = (int*)(rsoa + HFK_TASK_RUNS + 1);        //   first position,
    int* rstart = rx + HFK_TASK_RUNS + 1;              //   index of its first ket in the task's sequence,
    int* rlen = rstart + HFK_TASK_RUNS + 1;            //   segment length,
    int* rnpp = rlen + HFK_TASK_RUNS + 1;              //   primitive pairs per ket
    double* sbra = (double*)(rnpp + HFK_TASK_RUNS + 1);   // bra primitive-pair records, then their screen values
    for (int t = tid; t < NR * ldk; t += nthr) krow[t] = 0.0;
    const int nrun = task.z - task.y;              // <= HFK_TASK_RUNS (the planner's tasks)
    if (tid == 0) {
        int st = 0;
        for (int r = 0; r < nrun; ++r) {
            const KetRun q = a.runs[task.y + r];
            rjcd[r] = q.jcd;
            rsoa[r] = q.soa;
            rx[r] = q.x;
            rstart[r] = st;
            rlen[r] = q.len;
            rnpp[r] = q.npp;
            st += q.y - q.x;
        }
        rstart[nrun] = st;
    }
    const int A = a.pair_A[bp], B = a.pair_B[bp];
    int ia[NA], ib[NB];
    double sa[NA], sb[NB];
#pragma unroll
    for (int x = 0; x < NA; ++x) { ia[x] = a.sh_bfp[A * HFK_MAXCOMP + x]; sa[x] = a.sh_scale[A * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < NB; ++x) { ib[x] = a.sh_bfp[B * HFK_MAXCOMP + x]; sb[x] = a.sh_scale[B * HFK_MAXCOMP + x]; }
    const double* __restrict__ D = a.D;
    // the bra shells' density rows over the compact columns (the discarded
    // column reads the zero column n)
    if (drows) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double* Drow = D + (size_t)(r < NA ? ia[r] : ib[r - NA]) * ld;
            for (int col = tid; col < ldk; col += nthr) drow[r * ldk + col] = col < ncols ? Drow[a.collist[col]] : 0.0;
        }
    }
    // the task's bra primitive-pair records and their per-component screen
    // values, read by every lane in every ket iteration: staged once in LDS,
    // the K field premultiplied by 2 pi^(5/2) (the product the ERI code
    // forms first, so the rounding is the same)
    const int nbs = bprim1 - bprim0;
    double* sbq = sbra + nbs * PP_STRIDE;
    {
        const double* g = a.pp + (size_t)(a.pair_pp0[bp] + bprim0) * PP_STRIDE;
        for (int t = tid; t < nbs * PP_STRIDE; t += nthr) {
            const double v = g[t];
            sbra[t] = (t & (PP_STRIDE - 1)) == PP_K ? HFK_TWO_PI_2_5 * v : v;
        }
        if constexpr (MODE == 1) {
            const double* q = a.qcomp + (long long)g[PP_QOFF];
            for (int t = tid; t < nbs * NBRA; t += nthr) sbq[t] = q[t];
        }
    }
    __syncthreads();

    double Dab[NBRA];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) Dab[x * NB + y] = D[(size_t)ia[x] * ld + ib[y]];
    double jab[NBRA];
#pragma unroll
    for (int k = 0; k < NBRA; ++k) jab[k] = 0.0;
    double* __restrict__ Jcd = a.jcd;
    const int* __restrict__ kd = a.kdesc;
    const int nket = (int)a.nket;
    const bool kscl = a.kscale != nullptr;
    const double bqhi = a.pair_qhi[bp], bsch = a.pair_sch[bp];

    // lanes take the kets in run order (a wave meets one primitive-pair
    // count except at run boundaries); every workgroup runs the same number
    // of rounds, so the turn barriers are uniform
    const int len = rstart[nrun];
    // run holding the task's v-th ket: binary search over the staged starts
    auto run_of = [&](int v) {
        int lo = 0, hi = nrun - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rstart[mid] <= v) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    for (int base = 0; base < len; base += nthr) {
        const int it = base + tid;
        const bool in = it < len;
        int r = 0, pos = 0;
        if (in) {
            r = run_of(it);
            pos = rx[r] + (it - rstart[r]);
        }
        const bool act = in && quartet_selected_pos(a, bqhi, bsch, pos);
        // K' contributions: rows a (x < NA) and b, at columns c and d
        double kac[NA * NC], kad[NA * ND], kbc[NB * NC], kbd[NB * ND];
        int cc[NC], cd[ND];
        double jv[NKET];
#pragma unroll
        for (int k = 0; k < NKET; ++k) jv[k] = 0.0;
        if (act) {
            T Vt[NBRA * NKET];
            {
                const int slen = rlen[r];
                const KetSoA<NF> kr{a.ksoa + rsoa[r] + pos + (size_t)kprim0 * NF * slen, slen};
                eri_class<LA, LB, LC, LD, T, BOYS, MODE, 1>(sbra, sbq, nbs, kr, min(kprim1, rnpp[r]) - kprim0, a.bt, Vt,
                                                         a.qtau);
            }
            int ic[NC], id[ND];
            double sc[NC], sd[ND];
#pragma unroll
            for (int x = 0; x < NC; ++x) {
                ic[x] = kd[x * nket + pos];
                cc[x] = kd[(NC + ND + x) * nket + pos];
                sc[x] = kscl ? a.kscale[x * nket + pos] : 1.0;
            }
#pragma unroll
            for (int x = 0; x < ND; ++x) {
                id[x] = kd[(NC + x) * nket + pos];
                cd[x] = kd[(2 * NC + ND + x) * nket + pos];
                sd[x] = kscl ? a.kscale[(NC + x) * nket + pos] : 1.0;
            }
            const int flags = kd[2 * (NC + ND) * nket + pos];
            double f = 1.0;
            if (A == B) f *= 0.5;
            if (flags & 1) f *= 0.5;
            if (bp == a.ket_idx[pos]) f *= 0.5;
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
                jv[k] = 2.0 * s;
            }
            // density blocks D_bd, D_bc, D_ad, D_ac
            double Dbd[NB * ND], Dbc[NB * NC], Dad[NA * ND], Dac[NA * NC];
            if (drows) {
#pragma unroll
                for (int y = 0; y < NB; ++y) {
#pragma unroll
                    for (int w = 0; w < ND; ++w) Dbd[y * ND + w] = drow[(NA + y) * ldk + cd[w]];
#pragma unroll
                    for (int z = 0; z < NC; ++z) Dbc[y * NC + z] = drow[(NA + y) * ldk + cc[z]];
                }
#pragma unroll
                for (int x = 0; x < NA; ++x) {
#pragma unroll
                    for (int w = 0; w < ND; ++w) Dad[x * ND + w] = drow[x * ldk + cd[w]];
#pragma unroll
                    for (int z = 0; z < NC; ++z) Dac[x * NC + z] = drow[x * ldk + cc[z]];
                }
            } else {
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
            }
#pragma unroll
            for (int x = 0; x < NA; ++x) {
#pragma unroll
                for (int z = 0; z < NC; ++z) {
                    double s = 0.0;
#pragma unroll
                    for (int y = 0; y < NB; ++y)
#pragma unroll
                        for (int w = 0; w < ND; ++w) s += V[((x * NB + y) * NC + z) * ND + w] * Dbd[y * ND + w];
                    kac[x * NC + z] = s;
                }
#pragma unroll
                for (int w = 0; w < ND; ++w) {
                    double s = 0.0;
#pragma unroll
                    for (int y = 0; y < NB; ++y)
#pragma unroll
                        for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * Dbc[y * NC + z];
                    kad[x * ND + w] = s;
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
                    kbc[y * NC + z] = s;
                }
#pragma unroll
                for (int w = 0; w < ND; ++w) {
                    double s = 0.0;
#pragma unroll
                    for (int x = 0; x < NA; ++x)
#pragma unroll
                        for (int z = 0; z < NC; ++z) s += V[((x * NB + y) * NC + z) * ND + w] * Dac[x * NC + z];
                    kbd[y * ND + w] = s;
                }
            }
        }
        // this ket's J'_cd entry (zero when the quartet is skipped: every
        // entry is rewritten by every build)
        if (in) {
            const long long entry = rjcd[r] + (it - rstart[r]);
#pragma unroll
            for (int k = 0; k < NKET; ++k) Jcd[entry * NKET + k] = jv[k];
        }
        // K' rows in LDS.  With several waves, phase ph lets wave w add to the
        // rows r with r mod nwaves == (w + ph) mod nwaves: in every phase each
        // row takes adds from one wave only, and over the phases a row takes
        // them in a fixed wave order.
        for (int ph = 0; ph < nwaves; ++ph) {
            const int grp = (wave + ph) % nwaves;
            if (act) {
#pragma unroll
                for (int x = 0; x < NA; ++x) {
                    if (x % nwaves != grp) continue;
#pragma unroll
                    for (int z = 0; z < NC; ++z) atomic_add_lds(&krow[x * ldk + cc[z]], kac[x * NC + z]);
#pragma unroll
                    for (int v = 0; v < ND; ++v) atomic_add_lds(&krow[x * ldk + cd[v]], kad[x * ND + v]);
                }
#pragma unroll
                for (int y = 0; y < NB; ++y) {
                    if ((NA + y) % nwaves != grp) continue;
#pragma unroll
                    for (int z = 0; z < NC; ++z) atomic_add_lds(&krow[(NA + y) * ldk + cc[z]], kbc[y * NC + z]);
#pragma unroll
                    for (int v = 0; v < ND; ++v) atomic_add_lds(&krow[(NA + y) * ldk + cd[v]], kbd[y * ND + v]);
                }
            }
            if (nwaves > 1) __syncthreads();
        }
    }
    // J'_ab: wave reduction (a fixed butterfly), then the waves in order
#pragma unroll
    for (int b = 0; b < NBRA; ++b) {
        double v = jab[b];
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
        if (lane == 0) sred[wave * NBRA + b] = v;
    }
    __syncthreads();
    if (tid < NBRA) {
        double v = 0.0;
        for (int w = 0; w < nwaves; ++w) v += sred[w * NBRA + tid];
        a.jab[(size_t)a.jab_slot[tix] * NBRA + tid] = 2.0 * v;
    }
    // the staged K' rows to the task's slot (row by row: no index division)
    double* __restrict__ ks = a.kslot + (size_t)tix * NR * ncols;
#pragma unroll
    for (int r = 0; r < NR; ++r)
        for (int col = tid; col < ncols; col += nthr) ks[(size_t)r * ncols + col] = krow[r * ldk + col];
}

// 0, or -3 when the workgroup's LDS exceeds 64 KB
template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
int launch_tiles(const QuartetArgs& a, hipStream_t s)
{
    if (a.ntask <= 0) return 0;
    const size_t shm = hfk_jk_lds_bytes(LA, LB, a.ncols, a.bra_stage, a.drows);
    if (shm > 64 * 1024) return -3;
    long long done = 0;
    while (done < a.ntask) {   // grid.x limit
        const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
        QuartetArgs b = a;
        b.tasks = a.tasks + done;
        b.task0 = a.task0 + done;
        hipLaunchKernelGGL((k_jk_tiles<LA, LB, LC, LD, T, BOYS, MODE>), dim3((unsigned)cnt), dim3(a.block), shm, s, b);
        done += cnt;
    }
    return 0;
}

// op: HFK_OP_JK (mode 0 main, 1 correction), HFK_OP_FLAT, HFK_OP_DIAG.
template <int LA, int LB, int LC, int LD>
int dispatch(int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)
{
    if (op == HFK_OP_JK) {
#define HFK_T(T)                                                                        \
        if (boys == 0) {                                                                  \
            if (mode == 0) return launch_tiles<LA, LB, LC, LD, T, 0, 0>(a, s);            \
            return launch_tiles<LA, LB, LC, LD, T, 0, 1>(a, s);                           \
        }                                                                                 \
        if (mode == 0) return launch_tiles<LA, LB, LC, LD, T, 1, 0>(a, s);                \
        return launch_tiles<LA, LB, LC, LD, T, 1, 1>(a, s);
        if (prec == 0) { HFK_T(double) }
        HFK_T(float)
#undef HFK_T
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
