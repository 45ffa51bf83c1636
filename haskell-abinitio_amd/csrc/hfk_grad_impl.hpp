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
// VGPRs and AGPRs sets occupancy (512 per lane per SIMD); (ps|ps) and (pp|ss)
// need a little over 256, which alone would leave one wave per SIMD.
template <int LA, int LB, int LC, int LD>
constexpr int grad_min_waves()
{
    return (LA + LB + LC + LD == 2) ? 2 : 1;
}

template <int LA, int LB, int LC, int LD, int BOYS, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grad_min_waves<LA, LB, LC, LD>())))
k_grad_tiles(QuartetArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NKET = NC * ND, NBRA = NA * NB;
    extern __shared__ double gacc[];   // [natom][3] when grad_lds
    const int4 task = a.tasks[blockIdx.x];
    const int bp = task.x;
    const int bprim0 = (task.w >> 24) & 0xff, bprim1 = (task.w >> 16) & 0xff;
    const int kprim0 = (task.w >> 8) & 0xff, kprim1 = task.w & 0xff;
    const int ld = a.n + 1;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & (WAVE - 1);
    const bool lds = a.grad_lds != 0;
    double* __restrict__ acc = lds ? gacc : a.grad;
    if (lds)
        for (int t = tid; t < 3 * a.natom; t += nthr) gacc[t] = 0.0;
    __syncthreads();

    const int A = a.pair_A[bp], B = a.pair_B[bp];
    const int atA = a.sh_atom[A], atB = a.sh_atom[B];
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
    const double* bra = a.pp + (size_t)(a.pair_pp0[bp] + bprim0) * PP_STRIDE;
    const int nbp = bprim1 - bprim0;   // ranges are explicit: 0 <= b0 < b1 <= npp
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};

    const int len = task_ket_count(a, task);
    for (int it = tid; it < len; it += nthr) {
        const int kp = a.ket_idx[task_ket_pos(a, task, it)];
        if (!quartet_selected(a, bp, kp)) continue;
        const int C = a.pair_A[kp], Dd = a.pair_B[kp];
        int ic[NC], id[ND];
        double sc[NC], sd[ND];
#pragma unroll
        for (int x = 0; x < NC; ++x) { ic[x] = a.sh_bfp[C * HFK_MAXCOMP + x]; sc[x] = a.sh_scale[C * HFK_MAXCOMP + x]; }
#pragma unroll
        for (int x = 0; x < ND; ++x) { id[x] = a.sh_bfp[Dd * HFK_MAXCOMP + x]; sd[x] = a.sh_scale[Dd * HFK_MAXCOMP + x]; }
        double f = 4.0;
        if (A == B) f *= 0.5;
        if (C == Dd) f *= 0.5;
        if (bp == kp) f *= 0.5;
        double G[NBRA * NKET];
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int z = 0; z < NC; ++z)
#pragma unroll
                    for (int w = 0; w < ND; ++w) {
                        const double dcd = D[(size_t)ic[z] * ld + id[w]];
                        const double dac = D[(size_t)ia[x] * ld + ic[z]], dbd = D[(size_t)ib[y] * ld + id[w]];
                        const double dad = D[(size_t)ia[x] * ld + id[w]], dbc = D[(size_t)ib[y] * ld + ic[z]];
                        const double gam = Dab[x * NB + y] * dcd - 0.25 * (dac * dbd + dad * dbc);
                        G[((x * NB + y) * NC + z) * ND + w] = (f * sa[x] * sb[y] * sc[z] * sd[w]) * gam;
                    }
        const double* ket = a.pp + (size_t)(a.pair_pp0[kp] + kprim0) * PP_STRIDE;
        const int nkp = task_kprim_end(a, kprim1, kp) - kprim0;
        double g[9];
        grad_class<LA, LB, LC, LD, double, BOYS, MODE>(bra, nbp, ket, nkp, a.qcomp, a.bt, G, a.qtau, g);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gA[k] += g[k];
            gB[k] += g[3 + k];
        }
        const int atC = a.sh_atom[C], atD = a.sh_atom[Dd];
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

template <int LA, int LB, int LC, int LD, int BOYS, int MODE>
void launch_grad_tiles(const QuartetArgs& a, hipStream_t s)
{
    if (a.ntask <= 0) return;
    const size_t shm = a.grad_lds ? sizeof(double) * 3 * (size_t)a.natom : 0;
    long long done = 0;
    while (done < a.ntask) {
        const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
        QuartetArgs b = a;
        b.tasks = a.tasks + done;
        hipLaunchKernelGGL((k_grad_tiles<LA, LB, LC, LD, BOYS, MODE>), dim3((unsigned)cnt), dim3(a.block), shm, s, b);
        done += cnt;
    }
}

template <int LA, int LB, int LC, int LD>
int dispatch_grad(int boys, int mode, const QuartetArgs& a, hipStream_t s)
{
    if (boys == 0) {
        if (mode == 0) launch_grad_tiles<LA, LB, LC, LD, 0, 0>(a, s);
        else launch_grad_tiles<LA, LB, LC, LD, 0, 1>(a, s);
    } else {
        if (mode == 0) launch_grad_tiles<LA, LB, LC, LD, 1, 0>(a, s);
        else launch_grad_tiles<LA, LB, LC, LD, 1, 1>(a, s);
    }
    return 0;
}

}  // namespace
