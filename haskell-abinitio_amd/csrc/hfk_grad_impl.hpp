// Two-electron part of the RHF nuclear gradient.  Included by each generated
// class translation unit after grad_class<LA,LB,LC,LD,T,BOYS,MODE> (the
// derivative-integral contraction) and hfk_eri_impl.hpp.
//
// dE2/dR = 1/2 sum_{mu nu la si} d(mu nu|la si)/dR Gamma_{mu nu la si}, with
// Gamma = D_mn D_ls - 1/4 (D_ml D_ns + D_ms D_nl) for the total density D.
// Over canonical shell quartets carrying the permutation weight f (the J/K
// kernel's: 1/2 for A==B, for C==D and for equal pairs) the eight images
// give the per-quartet density 4 f Gamma.  The kernel runs on the J/K task
// lists (bra pair wave-uniform, ket pairs per lane, primitive ranges), so
// the same screen and the same quartets enter the energy and its gradient.
// A, B, C derivatives come from the contraction; D = -(A + B + C).
#pragma once

namespace {

// Waves per SIMD the register allocation must leave room for.  The total of
// VGPRs and AGPRs sets occupancy (512 per lane per SIMD); in FP64 (ps|ps) and
// (pp|ss) need a little over 256, which alone would leave one wave per SIMD.
// In FP32 (the FP32 part of a mixed-precision task list) (pp|ss) at three
// waves and (pp|ps) at two pay for a little spill; the other FP32 classes
// are faster at the occupancy their registers give.
template <int LA, int LB, int LC, int LD, typename T>
constexpr int grad_min_waves()
{
    constexpr int L = LA + LB + LC + LD;
    if constexpr (sizeof(T) == 4 && LA == 1 && LB == 1 && LC == 0) return 3;   // (pp|ss)
    if constexpr (sizeof(T) == 4 && LA == 1 && L == 3) return 2;               // (pp|ps)
    return L <= 1 ? 3 : (L == 2 ? 2 : 1);
}

// GRP: which of the kGradGroups blocks of bra components this launch
// contracts (a class whose density block does not fit one lane's registers
// splits its bra components over several launches; the partial sums add).
// T: the arithmetic of the derivative integrals (double; float for the FP32
// part of a mixed-precision task list); the density block is formed in FP64
// and the per-lane sums are kept in FP64.
template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE, int GRP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grad_min_waves<LA, LB, LC, LD, T>())))
k_grad_tiles(QuartetArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NKET = NC * ND, NBRA = NA * NB;
    constexpr int NSUB = NBRA / kGradGroups, SUB0 = GRP * NSUB;   // bra components [SUB0, SUB0 + NSUB)
    static_assert(NSUB * kGradGroups == NBRA, "bra components split evenly");
    constexpr int NR = NA + NB;
    constexpr int NF = hfk_ket_fields(LC, LD);
    extern __shared__ double gsm[];
    const int4 task = a.tasks[blockIdx.x];
    const int bp = task.x;
    const int bprim0 = (task.w >> 24) & 0xff, bprim1 = (task.w >> 16) & 0xff;
    const int kprim0 = (task.w >> 8) & 0xff, kprim1 = task.w & 0xff;
    const int ld = a.n + 1;
    const int ncols = a.ncols, ldk = ncols + 1;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & (WAVE - 1);
    const bool lds = a.grad_lds != 0, drows = a.drows != 0;
    // LDS: [natom][3] accumulator (grad_lds), [NR][ldk] density rows of the
    // bra shells (drows), the task's runs, the staged bra primitive pairs
    double* gacc = gsm;
    double* drow = gacc + (lds ? 3 * a.natom : 0);
    long long* rsoa = (long long*)(drow + (drows ? NR * ldk : 0));
    int* rx = (int*)(rsoa + HFK_TASK_RUNS + 1);
    int* rstart = rx + HFK_TASK_RUNS + 1;
    int* rlen = rstart + HFK_TASK_RUNS + 1;
    int* rnpp = rlen + HFK_TASK_RUNS + 1;
    double* sbra = (double*)(rnpp + HFK_TASK_RUNS + 1);
    double* __restrict__ acc = lds ? gacc : a.grad;
    if (lds)
        for (int t = tid; t < 3 * a.natom; t += nthr) gacc[t] = 0.0;
    const int nrun = task.z - task.y;
    if (tid == 0) {
        int st = 0;
        for (int r = 0; r < nrun; ++r) {
            const KetRun q = a.runs[task.y + r];
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
    const int atA = a.sh_atom[A], atB = a.sh_atom[B];
    int ia[NA], ib[NB];
    double sa[NA], sb[NB];
#pragma unroll
    for (int x = 0; x < NA; ++x) { ia[x] = a.sh_bfp[A * HFK_MAXCOMP + x]; sa[x] = a.sh_scale[A * HFK_MAXCOMP + x]; }
#pragma unroll
    for (int x = 0; x < NB; ++x) { ib[x] = a.sh_bfp[B * HFK_MAXCOMP + x]; sb[x] = a.sh_scale[B * HFK_MAXCOMP + x]; }
    const double* __restrict__ D = a.D;
    if (drows) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double* Drow = D + (size_t)(r < NA ? ia[r] : ib[r - NA]) * ld;
            for (int col = tid; col < ldk; col += nthr) drow[r * ldk + col] = col < ncols ? Drow[a.collist[col]] : 0.0;
        }
    }
    // the task's bra primitive-pair records and their per-component screen
    // values, read by every lane for every ket primitive: staged once, the K
    // field premultiplied by 2 pi^(5/2)
    const int nbp = bprim1 - bprim0;   // ranges are explicit: 0 <= b0 < b1 <= npp
    double* sbq = sbra + nbp * PP_STRIDE;
    {
        const double* gp = a.pp + (size_t)(a.pair_pp0[bp] + bprim0) * PP_STRIDE;
        for (int t = tid; t < nbp * PP_STRIDE; t += nthr) {
            const double v = gp[t];
            sbra[t] = (t & (PP_STRIDE - 1)) == PP_K ? HFK_TWO_PI_2_5 * v : v;
        }
        if constexpr (MODE == 1) {
            const double* q = a.qcomp + (long long)gp[PP_QOFF];
            for (int t = tid; t < nbp * NBRA; t += nthr) sbq[t] = q[t];
        }
    }
    __syncthreads();

    double Dab[NBRA];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) Dab[x * NB + y] = D[(size_t)ia[x] * ld + ib[y]];
    const int* __restrict__ kd = a.kdesc;
    const int nket = (int)a.nket;
    const bool kscl = a.kscale != nullptr;
    const double bqhi = a.pair_qhi[bp], bsch = a.pair_sch[bp];
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};

    const int len = rstart[nrun];
    for (int it = tid; it < len; it += nthr) {
        int lo = 0, hi = nrun - 1;   // run of the it-th ket
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rstart[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const int r = lo, pos = rx[r] + (it - rstart[r]);
        if (!quartet_selected_pos(a, bqhi, bsch, pos)) continue;
        int ic[NC], id[ND], cc[NC], cd[ND];
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
        double f = 4.0;
        if (A == B) f *= 0.5;
        if (flags & 1) f *= 0.5;
        if (bp == a.ket_idx[pos]) f *= 0.5;
        double G[NSUB * NKET];
#pragma unroll
        for (int si = 0; si < NSUB; ++si) {
            const int x = (SUB0 + si) / NB, y = (SUB0 + si) % NB;
#pragma unroll
                for (int z = 0; z < NC; ++z)
#pragma unroll
                    for (int w = 0; w < ND; ++w) {
                        const double dcd = D[(size_t)ic[z] * ld + id[w]];
                        double dac, dbd, dad, dbc;
                        if (drows) {
                            dac = drow[x * ldk + cc[z]];
                            dbd = drow[(NA + y) * ldk + cd[w]];
                            dad = drow[x * ldk + cd[w]];
                            dbc = drow[(NA + y) * ldk + cc[z]];
                        } else {
                            dac = D[(size_t)ia[x] * ld + ic[z]];
                            dbd = D[(size_t)ib[y] * ld + id[w]];
                            dad = D[(size_t)ia[x] * ld + id[w]];
                            dbc = D[(size_t)ib[y] * ld + ic[z]];
                        }
                        const double gam = Dab[x * NB + y] * dcd - 0.25 * (dac * dbd + dad * dbc);
                        G[(si * NC + z) * ND + w] = (f * sa[x] * sb[y] * sc[z] * sd[w]) * gam;
                    }
        }
        const int slen = rlen[r];
        const KetSoA<NF> kr{a.ksoa + rsoa[r] + pos + (size_t)kprim0 * NF * slen, slen};
        T Gt[NSUB * NKET], gt[9];
#pragma unroll
        for (int k = 0; k < NSUB * NKET; ++k) Gt[k] = (T)G[k];
        grad_class<LA, LB, LC, LD, T, BOYS, MODE, GRP>(sbra, sbq, nbp, kr, min(kprim1, rnpp[r]) - kprim0, a.bt, Gt, a.qtau,
                                                       gt);
        double g[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = (double)gt[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gA[k] += g[k];
            gB[k] += g[3 + k];
        }
        const int atC = kd[(2 * (NC + ND) + 1) * nket + pos], atD = kd[(2 * (NC + ND) + 2) * nket + pos];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            atomic_add_f64(&acc[3 * atC + k], g[6 + k]);
            atomic_add_f64(&acc[3 * atD + k], -((g[k] + g[3 + k]) + g[6 + k]));
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double u = gA[k], v = gB[k];
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            u += __shfl_xor(u, off, WAVE);
            v += __shfl_xor(v, off, WAVE);
        }
        if (lane == 0) {
            atomic_add_f64(&acc[3 * atA + k], u);
            atomic_add_f64(&acc[3 * atB + k], v);
        }
    }
    if (lds) {
        __syncthreads();
        for (int t = tid; t < 3 * a.natom; t += nthr) {
            const double v = gacc[t];
            if (v != 0.0) atomic_add_f64(&a.grad[t], v);
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE, int GRP>
void launch_grad_group(const QuartetArgs& a, hipStream_t s)
{
    const size_t shm = hfk_grad_lds_bytes(LA, LB, a.ncols, a.grad_lds ? a.natom : 0, a.drows, a.bra_stage);
    long long done = 0;
    while (done < a.ntask) {
        const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
        QuartetArgs b = a;
        b.tasks = a.tasks + done;
        hipLaunchKernelGGL((k_grad_tiles<LA, LB, LC, LD, T, BOYS, MODE, GRP>), dim3((unsigned)cnt), dim3(a.block), shm, s, b);
        done += cnt;
    }
    if constexpr (GRP + 1 < kGradGroups) launch_grad_group<LA, LB, LC, LD, T, BOYS, MODE, GRP + 1>(a, s);
}

template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE>
void launch_grad_tiles(const QuartetArgs& a, hipStream_t s)
{
    if (a.ntask <= 0) return;
    launch_grad_group<LA, LB, LC, LD, T, BOYS, MODE, 0>(a, s);
}

// prec: 0 FP64, 1 FP32 (the FP32 part of a mixed-precision task list)
template <int LA, int LB, int LC, int LD>
int dispatch_grad(int boys, int mode, int prec, const QuartetArgs& a, hipStream_t s)
{
#define HFK_G(T)                                                                     \
    if (boys == 0) {                                                                 \
        if (mode == 0) launch_grad_tiles<LA, LB, LC, LD, T, 0, 0>(a, s);             \
        else launch_grad_tiles<LA, LB, LC, LD, T, 0, 1>(a, s);                       \
    } else {                                                                         \
        if (mode == 0) launch_grad_tiles<LA, LB, LC, LD, T, 1, 0>(a, s);             \
        else launch_grad_tiles<LA, LB, LC, LD, T, 1, 1>(a, s);                       \
    }
    if (prec == 0) { HFK_G(double) }
    else { HFK_G(float) }
#undef HFK_G
    return 0;
}

}  // namespace
