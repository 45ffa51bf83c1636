// Device-side planning of the quartet launches: the pair classes in their
// screen order, the staircase of surviving kets, the task and run lists of
// the J/K and gradient kernels (with their precision parts), the
// costliest-first task order and the reduction lists of the J/K slots.
// Everything that grows with the number of pairs or tasks is built by
// kernels (radix sorts and scans from hipCUB); the host keeps the few
// numbers per launch that size the plan (segment boundaries of each class,
// primitive-range steps, totals).
//
// Orders (the kernels and the fixed-order reductions rely on them):
//  - class lists: pair class, then primitive-pair count ascending, then
//    largest screen value (QHI) descending, then pair index; a class's runs
//    of equal primitive-pair count are its segments;
//  - tasks: by part (FP64, FP32), within a part by bra position, for each
//    bra its multi-run tasks then its primitive-split tasks (segment,
//    primitive piece, ket chunk), then stably costliest first;
//  - runs: in the order the tasks create them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "hfk_host.hpp"

namespace {

// device temporaries of one planning call, freed at its end
// Temporaries of one planning call, bump-allocated (256-byte aligned) from
// chunks of at least 8 MB: a small molecule's plan makes a few allocations
// instead of one hipMalloc/hipFree pair per array.  Freed together.
struct Tmp {
    static constexpr size_t CHUNK = 8u << 20, ALIGN = 256;
    std::vector<void*> chunks;
    char* cur = nullptr;
    size_t left = 0;
    ~Tmp()
    {
        for (void* q : chunks) hipFree(q);
    }
    template <typename T>
    int alloc(T** x, size_t n)
    {
        const size_t bytes = ((std::max<size_t>(n, 1) * sizeof(T) + ALIGN - 1) / ALIGN) * ALIGN;
        if (bytes > left) {
            const size_t sz = std::max(bytes, CHUNK);
            void* q = nullptr;
            hipError_t e = hipMalloc(&q, sz);
            if (e != hipSuccess) return hfk_set_error(HFK_ENOMEM, "planning: hipMalloc(%zu bytes) failed", sz);
            chunks.push_back(q);
            cur = (char*)q;
            left = sz;
        }
        *x = (T*)cur;
        cur += bytes;
        left -= bytes;
        return HFK_OK;
    }
};

int nc_of(int l) { return (l + 1) * (l + 2) / 2; }

unsigned grid_of(long long n, int block)
{
    long long g = (n + block - 1) / block;
    return (unsigned)std::max(1LL, g);
}

// stable radix sort of (key, value) pairs on the low end_bit bits of the keys
template <typename K, typename V>
int sort_pairs(hfk_ctx* c, Tmp& t, const K* kin, K* kout, const V* vin, V* vout, long long n, int end_bit)
{
    if (n <= 0) return HFK_OK;
    size_t bytes = 0;
    HFK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, c->stream));
    char* w = nullptr;
    HFK_TRY(t.alloc(&w, bytes));
    HFK_HIP(hipcub::DeviceRadixSort::SortPairs((void*)w, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, c->stream));
    return HFK_OK;
}

// out[0..n) = exclusive prefix sums of in[0..n); out[n] = the total
template <typename T>
int scan_total(hfk_ctx* c, Tmp& t, const T* in, T* out, long long n)
{
    HFK_HIP(hipMemsetAsync(out, 0, sizeof(T), c->stream));
    if (n <= 0) return HFK_OK;
    size_t bytes = 0;
    HFK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out + 1, (int)n, c->stream));
    char* w = nullptr;
    HFK_TRY(t.alloc(&w, bytes));
    HFK_HIP(hipcub::DeviceScan::InclusiveSum((void*)w, bytes, in, out + 1, (int)n, c->stream));
    return HFK_OK;
}

template <typename T>
int to_dev(hfk_ctx* c, Tmp& t, T** d, const std::vector<T>& h)
{
    HFK_TRY(t.alloc(d, h.size()));
    if (!h.empty()) HFK_HIP(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
    return HFK_OK;
}

// one block-wide sum of a 64-bit value per block, added atomically (exact
// integer sums: the order does not matter)
__device__ void block_add(unsigned long long v, unsigned long long* out)
{
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(out, red[0]);
}

// ----------------------------------------------------------------------
// pair classes
// ----------------------------------------------------------------------

// Sort keys of every pair: the largest screen value descending (the bits of
// a non-negative double order like the value), and class << 8 | primitive
// pairs, class HFK_NPC for pairs the pair screen drops.
__global__ void k_class_keys(int npair, const int* pair_A, const int* pair_B, const int* sh_L, const int* pair_npp,
                             const double* qhi, const double* gmax, double tau, int ref, unsigned long long* kq,
                             unsigned* kc, int* idx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const double q = qhi[p];   // >= 0 (NaN screen values are +inf here)
    kq[p] = ~(unsigned long long)__double_as_longlong(q);
    const bool keep = !(ref && q * gmax[0] < tau);
    const int la = sh_L[pair_A[p]], lb = sh_L[pair_B[p]];
    const int cls = la * (la + 1) / 2 + lb;   // la >= lb by construction
    kc[p] = ((unsigned)(keep ? cls : HFK_NPC) << 8) | (unsigned)pair_npp[p];
    idx[p] = p;
}

__global__ void k_gather_u32(int n, const int* idx, const unsigned* src, unsigned* dst)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

// counts per class and per (class, primitive-pair count)
__global__ void k_class_hist(int n, const unsigned* kc, int* cls_cnt, int* npp_hist)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const unsigned k = kc[j] >> 8, m = kc[j] & 0xff;
    atomicAdd(&cls_cnt[k], 1);
    if (k < HFK_NPC) atomicAdd(&npp_hist[k * 256 + m], 1);
}

__global__ void k_sqrt(int n, const double* in, double* out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = sqrt(in[j]);
}

// ----------------------------------------------------------------------
// plan of one launch
// ----------------------------------------------------------------------

__global__ void k_gather_d(int n, const int* idx, const double* src, double* dst)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

__global__ void k_gather_i(int n, const int* idx, const int* src, int* dst)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

// segment of a class position: the last s with first[s] <= j
__device__ __forceinline__ int seg_of(const int* first, int nseg, int j)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Staircase: the kets of bra i in ket segment s that can pass the reference
// screen are [first_s, jend): the segment is in descending QHI order and a
// ket needs QHI >= tau / QHI_i (relaxed by 1e-12 so that rounding never
// drops a survivor); a class paired with itself keeps positions <= i.  Also
// the estimated primitive quartets of the launch.
__global__ void k_staircase(int nb, int nkseg, const double* bq, const double* kq, const int* kfirst, const int* knpp,
                            const int* bnpp, int same, int ref, double tau, int* jend, unsigned long long* est)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long e = 0;
    if (t < (long long)nb * nkseg) {
        const int i = (int)(t / nkseg), s = (int)(t - (long long)i * nkseg);
        const double qi = bq[i];
        const double need = (!ref || isinf(qi)) ? -1.0 : (qi > 0.0 ? tau / qi * (1.0 - 1e-12) : INFINITY);
        const long long first = kfirst[s];
        long long hi_end = kfirst[s + 1];
        if (same) hi_end = min(hi_end, (long long)i + 1);
        long long end = first;
        if (hi_end > first) {
            long long lo = first, hi = hi_end;
            while (lo < hi) {
                const long long mid = (lo + hi) / 2;
                if (kq[mid] >= need) lo = mid + 1;
                else hi = mid;
            }
            end = lo;
        }
        jend[t] = (int)end;
        e = (unsigned long long)bnpp[i] * (unsigned long long)knpp[s] * (unsigned long long)(end - first);
    }
    block_add(e, est);
}

// suffix maxima of the kets' Schwarz bounds within each segment
__global__ void k_suffix_max(int nkseg, const int* kfirst, const double* ksch, double* smax)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nkseg) return;
    double m = 0.0;
    for (int j = kfirst[s + 1] - 1; j >= kfirst[s]; --j) {
        m = fmax(ksch[j], m);
        smax[j] = m;
    }
}

// precision split: the first ket from which every bound times the bra's is
// below the threshold (FP32 from there on)
__global__ void k_split(int nb, int nkseg, const double* bsch, const double* smax, const int* kfirst, const int* jend,
                        double thr, int* jmid)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * nkseg) return;
    const int i = (int)(t / nkseg), s = (int)(t - (long long)i * nkseg);
    const double bs = bsch[i];
    long long lo = kfirst[s], hi = jend[t];
    while (lo < hi) {
        const long long mid = (lo + hi) / 2;
        if (bs * smax[mid] < thr) hi = mid;
        else lo = mid + 1;
    }
    jmid[t] = (int)lo;
}

__device__ __forceinline__ bool whole_pair(int nbp, int nkp, int single, int per_lane)
{
    return !single && nbp <= HFK_BRA_STAGE && nbp * nkp <= per_lane;
}

// kets of the segments that join the multi-run tasks
__global__ void k_merged(int nb, int nkseg, const int* jend, const int* kfirst, const int* knpp, const int* bnpp,
                         int single, int per_lane, unsigned long long* out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (t < (long long)nb * nkseg) {
        const int i = (int)(t / nkseg), s = (int)(t - (long long)i * nkseg);
        if (whole_pair(bnpp[i], knpp[s], single, per_lane)) v = (unsigned long long)(jend[t] - kfirst[s]);
    }
    block_add(v, out);
}

// J'_cd entries of (bra i, ket segment s): pieces x surviving kets
__global__ void k_jcd_count(int nb, int nkseg, int nbseg, const int* bfirst, const int* npiece, const int* jend,
                            const int* kfirst, long long* v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * nkseg) return;
    const int i = (int)(t / nkseg), s = (int)(t - (long long)i * nkseg);
    const int b = seg_of(bfirst, nbseg, i);
    v[t] = (long long)npiece[(size_t)b * nkseg + s] * (jend[t] - kfirst[s]);
}

__global__ void k_add_offset(int n, const long long* in, const long long* base, long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] + base[0];
}

struct WalkArgs {
    int nb, nkseg, part, split, single, coop, per_lane, chunk;
    long long cmerge;
    const int* bl;              // bra pairs by position
    const int* bnpp;            // their primitive-pair counts
    const int* kfirst;          // [nkseg + 1]
    const int* knpp;            // [nkseg]
    const long long* soa_base;  // [nkseg]
    const int* jend;            // [nb][nkseg]
    const int* jmid;            // [nb][nkseg] (split)
    const long long* jbase;     // [nb][nkseg]
    // count pass
    long long* ntask;           // [nb]
    long long* nrun;            // [nb]
    // write pass
    const long long* toff;      // [nb] first task of bra i (launch numbering)
    const long long* roff;      // [nb] first run
    int4* tasks;
    KetRun* runs;
    long long* cost;
    int* stage;                 // max bra primitive range staged (atomicMax)
    int* lenmax;                // longest task (atomicMax)
};

// The tasks and runs of bra i in one precision part, as the plan defines
// them (see the file comment); counts only, or writes them.
template <bool WRITE>
__global__ void k_task_walk(WalkArgs w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nb) return;
    const int nbp = w.bnpp[i], bp = w.bl[i];
    const size_t ib = (size_t)i * w.nkseg;
    const long long tb = WRITE ? w.toff[i] : 0, rb = WRITE ? w.roff[i] : 0;
    long long nt = 0, nr = 0, r0 = 0, cacc = 0;
    int len = 0, stage = 0, lmax = 0;
    // multi-run tasks: the kets of the unsplit segments, cut every cmerge
    // kets or HFK_TASK_RUNS runs
    for (int s = 0; s < w.nkseg; ++s) {
        const int nkp = w.knpp[s];
        if (!whole_pair(nbp, nkp, w.single, w.per_lane)) continue;
        long long x = w.part == 0 ? w.kfirst[s] : w.jmid[ib + s];
        const long long x1 = (w.part == 0 && w.split) ? w.jmid[ib + s] : w.jend[ib + s];
        while (x < x1) {
            const long long take = min(x1 - x, w.cmerge - len);
            if (WRITE) {
                KetRun r;
                r.x = (int)x;
                r.y = (int)(x + take);
                r.jcd = w.jbase[ib + s] + (x - w.kfirst[s]);
                r.soa = w.soa_base[s] - w.kfirst[s];
                r.len = w.kfirst[s + 1] - w.kfirst[s];
                r.npp = nkp;
                w.runs[rb + nr] = r;
            }
            ++nr;
            cacc += take * nbp * nkp;
            x += take;
            len += (int)take;
            if (len == w.cmerge || nr - r0 == HFK_TASK_RUNS) {
                if (WRITE) {
                    w.tasks[tb + nt] = make_int4(bp, (int)(rb + r0), (int)(rb + nr), (nbp << 16) | 0xff);
                    w.cost[tb + nt] = cacc;
                }
                ++nt;
                cacc = 0;
                lmax = max(lmax, len);
                r0 = nr;
                len = 0;
            }
        }
    }
    if (len > 0) {
        if (WRITE) {
            w.tasks[tb + nt] = make_int4(bp, (int)(rb + r0), (int)(rb + nr), (nbp << 16) | 0xff);
            w.cost[tb + nt] = cacc;
        }
        ++nt;
        cacc = 0;
        lmax = max(lmax, len);
    }
    if (nbp > 0) stage = w.coop ? nbp : min(nbp, HFK_BRA_STAGE);
    // segments split into primitive ranges: one run per task
    for (int s = 0; s < w.nkseg; ++s) {
        const int nkp = w.knpp[s];
        if (whole_pair(nbp, nkp, w.single, w.per_lane)) continue;
        const long long k0 = w.kfirst[s], k1 = w.jend[ib + s];
        const long long p0 = w.part == 0 ? w.kfirst[s] : w.jmid[ib + s];
        const long long p1 = (w.part == 0 && w.split) ? w.jmid[ib + s] : w.jend[ib + s];
        if (p1 <= p0) continue;
        int bstep = nbp, kstep = nkp;
        if (!w.coop && nbp * nkp > w.per_lane) {
            bstep = max(1, min(nbp, w.per_lane));
            kstep = max(1, min(nkp, w.per_lane / bstep));
        }
        if (!w.coop) bstep = min(bstep, HFK_BRA_STAGE);
        int piece = 0;
        for (int b0 = 0; b0 < nbp; b0 += bstep) {
            const int b1 = min(nbp, b0 + bstep);
            for (int q0 = 0; q0 < nkp; q0 += kstep, ++piece) {
                const int q1 = min(nkp, q0 + kstep);
                const int wbits = (b0 << 24) | (b1 << 16) | (q0 << 8) | q1;
                for (long long j0 = p0; j0 < p1; j0 += w.chunk) {
                    const long long j1 = min(p1, j0 + (long long)w.chunk);
                    if (WRITE) {
                        KetRun r;
                        r.x = (int)j0;
                        r.y = (int)j1;
                        r.jcd = w.jbase[ib + s] + (long long)piece * (k1 - k0) + (j0 - k0);
                        r.soa = w.soa_base[s] - w.kfirst[s];
                        r.len = w.kfirst[s + 1] - w.kfirst[s];
                        r.npp = nkp;
                        w.runs[rb + nr] = r;
                        w.tasks[tb + nt] = make_int4(bp, (int)(rb + nr), (int)(rb + nr + 1), wbits);
                        w.cost[tb + nt] = (j1 - j0) * (long long)(b1 - b0) * (q1 - q0);
                    }
                    ++nr;
                    ++nt;
                    lmax = max(lmax, (int)(j1 - j0));
                }
                stage = max(stage, b1 - b0);
            }
        }
    }
    if (WRITE) {
        atomicMax(w.stage, stage);
        atomicMax(w.lenmax, lmax);
    } else {
        w.ntask[i] = nt;
        w.nrun[i] = nr;
    }
}

// costliest first: 255 - the quarter-octave of the estimated cost (4 log2
// cost, from its two bits below the leading one)
__global__ void k_cost_keys(long long n, const long long* cost, unsigned char* key, int* idx)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long v = cost[t];
    int bucket = 0;
    if (v > 0) {
        const int b = 63 - __clzll((unsigned long long)v);
        const int sub = b >= 2 ? (int)((v >> (b - 2)) & 3) : (int)((v << (2 - b)) & 3);
        bucket = 4 * b + sub;
    }
    key[t] = (unsigned char)(255 - bucket);
    idx[t] = (int)t;
}

__global__ void k_gather_tasks(long long n, const int* idx, const int4* src, int4* dst)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[idx[t]];
}

// reduction lists: (shell, entry) keys for the K' rows of the tasks' bra
// shells (A then B per task), and (bra position, task) for the J'_ab slots
__global__ void k_list_keys(long long ntask, const int4* tasks, const int* pair_A, const int* pair_B, const int* pos_of,
                            unsigned* kkey, int* kval, unsigned* bkey, int* bval, int* kcnt, int* bcnt)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    const int bp = tasks[t].x;
    const int A = pair_A[bp], B = pair_B[bp], pos = pos_of[bp];
    kkey[2 * t] = (unsigned)A;
    kval[2 * t] = (int)(2 * t);
    kkey[2 * t + 1] = (unsigned)B;
    kval[2 * t + 1] = (int)(2 * t + 1);
    bkey[t] = (unsigned)pos;
    bval[t] = (int)t;
    atomicAdd(&kcnt[A], 1);
    atomicAdd(&kcnt[B], 1);
    atomicAdd(&bcnt[pos], 1);
}

__global__ void k_list_out(long long ntask, int na, const int* kval, const int* bval, int2* kcsr, int* jab_slot)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 2 * ntask) {
        const int v = kval[j];
        kcsr[j] = make_int2(v >> 1, (v & 1) ? na : 0);
    }
    if (j < ntask) jab_slot[bval[j]] = (int)j;
}

__global__ void k_pos_of(int nb, const int* bl, int* pos_of)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) pos_of[bl[i]] = i;
}

// K reduction jobs per shell component row: HFK_KSEG entries each
__global__ void k_kjob_count(int nshell, const int* koff, const int* sh_bf, int* njob, int* nrow)
{
    const int S = blockIdx.x * blockDim.x + threadIdx.x;
    if (S >= nshell) return;
    const int q0 = koff[S], q1 = koff[S + 1];
    int nj = 0, nrw = 0;
    if (q1 > q0)
        for (int x = 0; x < HFK_MAXCOMP; ++x)
            if (sh_bf[S * HFK_MAXCOMP + x] >= 0) {
                nj += (q1 - q0 + HFK_KSEG - 1) / HFK_KSEG;
                ++nrw;
            }
    njob[S] = nj;
    nrow[S] = nrw;
}

__global__ void k_kjob_write(int nshell, const int* koff, const int* sh_bf, const int* joff, const int* roff, int4* kjob,
                             int4* krow)
{
    const int S = blockIdx.x * blockDim.x + threadIdx.x;
    if (S >= nshell) return;
    const int q0 = koff[S], q1 = koff[S + 1];
    if (q1 <= q0) return;
    int j = joff[S], r = roff[S];
    for (int x = 0; x < HFK_MAXCOMP; ++x) {
        if (sh_bf[S * HFK_MAXCOMP + x] < 0) continue;
        const int j0 = j;
        for (int q = q0; q < q1; q += HFK_KSEG) kjob[j++] = make_int4(S * HFK_MAXCOMP + x, q, min(q1, q + HFK_KSEG), 0);
        krow[r++] = make_int4(S * HFK_MAXCOMP + x, j0, j - j0, 0);
    }
}

// per ket position: descriptor words, screen values, component scales, and
// its place in the position-major records and in the J'_cd layout
__global__ void k_ket_desc(int nk, int nc, int nd, const int* kl, const int* pair_A, const int* pair_B,
                           const int* sh_bfp, const double* sh_scale, const int* sh_atom, const int* colmap,
                           const double* qhi, const double* sch, const int* kfirst, int nkseg, const long long* soa_base,
                           const int* knpp, int* kdesc, double* kposq, double* kscale, int* nonunit,
                           long long* pos_soa, int* pos_len, int* pos_npp, int* pos_kseg)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nk) return;
    const int p = kl[j], C = pair_A[p], D = pair_B[p];
    int nu = 0;
    for (int x = 0; x < nc; ++x) {
        const int f = sh_bfp[C * HFK_MAXCOMP + x];
        kdesc[(size_t)x * nk + j] = f;
        kdesc[(size_t)(nc + nd + x) * nk + j] = colmap[f];
        const double sc = sh_scale[C * HFK_MAXCOMP + x];
        kscale[(size_t)x * nk + j] = sc;
        nu |= sc != 1.0;
    }
    for (int x = 0; x < nd; ++x) {
        const int f = sh_bfp[D * HFK_MAXCOMP + x];
        kdesc[(size_t)(nc + x) * nk + j] = f;
        kdesc[(size_t)(2 * nc + nd + x) * nk + j] = colmap[f];
        const double sc = sh_scale[D * HFK_MAXCOMP + x];
        kscale[(size_t)(nc + x) * nk + j] = sc;
        nu |= sc != 1.0;
    }
    kdesc[(size_t)(2 * (nc + nd)) * nk + j] = C == D ? 1 : 0;
    kdesc[(size_t)(2 * (nc + nd) + 1) * nk + j] = sh_atom[C];
    kdesc[(size_t)(2 * (nc + nd) + 2) * nk + j] = sh_atom[D];
    kposq[j] = qhi[p];
    kposq[nk + j] = sch[p];
    if (nu) atomicOr(nonunit, 1);
    const int s = seg_of(kfirst, nkseg, j);
    pos_soa[j] = soa_base[s] - kfirst[s];
    pos_len[j] = kfirst[s + 1] - kfirst[s];
    pos_npp[j] = knpp[s];
    pos_kseg[j] = s;
}

}  // namespace

// ----------------------------------------------------------------------
// Pair classes on the device: the pair screen (a pair whose largest screen
// value times the largest over all pairs is below tau can never pass the
// primitive screen), then two stable radix sorts (QHI descending, then class
// and primitive-pair count) over the pair indices; the class lists are the
// runs of one class, their segments the runs of one primitive-pair count.
// ----------------------------------------------------------------------
int hfk_ctx_classify_pairs(hfk_ctx* c)
{
    Tmp t;
    const int np = c->npair;
    const bool ref = c->opt.screen_mode == HFK_SCREEN_REFERENCE && c->opt.screen_tau > 0.0;
    double* d_gmax = nullptr;
    HFK_TRY(t.alloc(&d_gmax, 1));
    {
        size_t bytes = 0;
        HFK_HIP(hipcub::DeviceReduce::Max(nullptr, bytes, c->d_pair_qhi, d_gmax, np, c->stream));
        char* w = nullptr;
        HFK_TRY(t.alloc(&w, bytes));
        HFK_HIP(hipcub::DeviceReduce::Max((void*)w, bytes, c->d_pair_qhi, d_gmax, np, c->stream));
    }
    unsigned long long *kq = nullptr, *kq2 = nullptr;
    unsigned *kc = nullptr, *kc2 = nullptr, *kc3 = nullptr;
    int *idx = nullptr, *idx2 = nullptr, *idx3 = nullptr, *cls_cnt = nullptr, *npp_hist = nullptr;
    HFK_TRY(t.alloc(&kq, np));
    HFK_TRY(t.alloc(&kq2, np));
    HFK_TRY(t.alloc(&kc, np));
    HFK_TRY(t.alloc(&kc2, np));
    HFK_TRY(t.alloc(&kc3, np));
    HFK_TRY(t.alloc(&idx, np));
    HFK_TRY(t.alloc(&idx2, np));
    HFK_TRY(t.alloc(&idx3, np));
    HFK_TRY(t.alloc(&cls_cnt, HFK_NPC + 1));
    HFK_TRY(t.alloc(&npp_hist, HFK_NPC * 256));
    HFK_HIP(hipMemsetAsync(cls_cnt, 0, sizeof(int) * (HFK_NPC + 1), c->stream));
    HFK_HIP(hipMemsetAsync(npp_hist, 0, sizeof(int) * HFK_NPC * 256, c->stream));
    hipLaunchKernelGGL(k_class_keys, dim3(grid_of(np, 256)), dim3(256), 0, c->stream, np, c->d_pair_A, c->d_pair_B,
                       c->d_sh_L, c->d_pair_npp, c->d_pair_qhi, d_gmax, c->opt.screen_tau, ref ? 1 : 0, kq, kc, idx);
    HFK_TRY(sort_pairs(c, t, kq, kq2, idx, idx2, np, 64));
    hipLaunchKernelGGL(k_gather_u32, dim3(grid_of(np, 256)), dim3(256), 0, c->stream, np, idx2, kc, kc2);
    HFK_TRY(sort_pairs(c, t, kc2, kc3, idx2, idx3, np, 16));
    hipLaunchKernelGGL(k_class_hist, dim3(grid_of(np, 256)), dim3(256), 0, c->stream, np, kc3, cls_cnt, npp_hist);
    HFK_HIP(hipGetLastError());
    std::vector<int> cnt(HFK_NPC + 1), hist((size_t)HFK_NPC * 256);
    HFK_HIP(hipMemcpyAsync(cnt.data(), cls_cnt, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipMemcpyAsync(hist.data(), npp_hist, sizeof(int) * hist.size(), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    long long off = 0, kept = 0;
    for (int k = 0; k < HFK_NPC; ++k) {
        if (c->d_class[k]) { dev_free(c, c->d_class[k]); c->d_class[k] = nullptr; }
        if (c->d_ksoa[k]) { dev_free(c, c->d_ksoa[k]); c->d_ksoa[k] = nullptr; }
        c->class_count[k] = cnt[k];
        c->class_segs[k].clear();
        long long first = 0;
        for (int m = 0; m < 256; ++m) {
            const int h = hist[(size_t)k * 256 + m];
            if (h == 0) continue;
            c->class_segs[k].push_back({first, first + h, m});
            first += h;
        }
        if (cnt[k] > 0) {
            HFK_TRY(dev_alloc(&c->d_class[k], (size_t)cnt[k]));
            HFK_HIP(hipMemcpyAsync(c->d_class[k], idx3 + off, sizeof(int) * cnt[k], hipMemcpyDeviceToDevice, c->stream));
        }
        off += cnt[k];
        kept += cnt[k];
    }
    HFK_HIP(hipStreamSynchronize(c->stream));
    c->stats.nshell_pair = kept;
    return HFK_OK;
}

// True Schwarz bounds sqrt(max_ab |(ab|ab)|) from the diagonal, on the device.
void hfk_launch_sqrt(int n, const double* in, double* out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_sqrt, dim3(grid_of(n, 256)), dim3(256), 0, s, n, in, out);
}

// ----------------------------------------------------------------------
// Task lists and reduction lists of one plan entry (device), see the file
// comment.  Fills the entry's device arrays and scalars (e.ncols and
// e.d_colmap set by the caller); the slots are allocated by the caller.
// ----------------------------------------------------------------------
int hfk_plan_entry_device(hfk_ctx* c, PlanEntry& e)
{
    Tmp t;
    const bool ref = c->opt.screen_mode == HFK_SCREEN_REFERENCE && c->opt.screen_tau > 0.0;
    const bool coop = hfk_class_coop(e.la, e.lb, e.lc, e.ld);
    const int block = hfk_plan_block(c, e);
    const int NA = nc_of(e.la), NC = nc_of(e.lc), ND = nc_of(e.ld);
    const int nbra_comp = NA * nc_of(e.lb), nket_comp = NC * ND;
    const int coop_s = std::max(1, 64 / nbra_comp);
    // kets per task of the primitive-split (heavy) segments: every task
    // stores its K' rows, so large systems take longer ones
    const int chunk = coop ? 16 * coop_s : block * c->kets_per_lane * (c->n >= 200 ? 16 : 1);
    const int nb = (int)c->class_count[e.kb], nk = (int)c->class_count[e.kk];
    const int* d_bl = c->d_class[e.kb];
    const int* d_kl = c->d_class[e.kk];
    const std::vector<ClassSeg>& ksegs = c->class_segs[e.kk];
    const std::vector<ClassSeg>& bsegs = c->class_segs[e.kb];
    const int nkseg = (int)ksegs.size(), nbseg = (int)bsegs.size();
    const int NF = hfk_ket_fields(e.lc, e.ld);
    // segment tables (host, a few entries)
    std::vector<int> kfirst(nkseg + 1), knpp(nkseg), bfirst(nbseg + 1), bnpp_seg(nbseg);
    std::vector<long long> soa_base(nkseg + 1, 0);
    for (int s = 0; s < nkseg; ++s) {
        kfirst[s] = (int)ksegs[s].first;
        knpp[s] = ksegs[s].npp;
        soa_base[s + 1] = soa_base[s] + (long long)NF * knpp[s] * (ksegs[s].end - ksegs[s].first);
    }
    kfirst[nkseg] = nk;
    for (int b = 0; b < nbseg; ++b) {
        bfirst[b] = (int)bsegs[b].first;
        bnpp_seg[b] = bsegs[b].npp;
    }
    bfirst[nbseg] = nb;
    int *d_kfirst, *d_knpp, *d_bfirst;
    long long* d_soa;
    HFK_TRY(to_dev(c, t, &d_kfirst, kfirst));
    HFK_TRY(to_dev(c, t, &d_knpp, knpp));
    HFK_TRY(to_dev(c, t, &d_bfirst, bfirst));
    HFK_TRY(to_dev(c, t, &d_soa, soa_base));
    // screen values, Schwarz bounds and primitive-pair counts by position
    double *bq, *kq, *bsch, *ksch;
    int* bnpp;
    HFK_TRY(t.alloc(&bq, nb));
    HFK_TRY(t.alloc(&kq, nk));
    HFK_TRY(t.alloc(&bsch, nb));
    HFK_TRY(t.alloc(&ksch, nk));
    HFK_TRY(t.alloc(&bnpp, nb));
    hipLaunchKernelGGL(k_gather_d, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, d_bl, c->d_pair_qhi, bq);
    hipLaunchKernelGGL(k_gather_d, dim3(grid_of(nk, 256)), dim3(256), 0, c->stream, nk, d_kl, c->d_pair_qhi, kq);
    hipLaunchKernelGGL(k_gather_d, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, d_bl, c->d_pair_sch, bsch);
    hipLaunchKernelGGL(k_gather_d, dim3(grid_of(nk, 256)), dim3(256), 0, c->stream, nk, d_kl, c->d_pair_sch, ksch);
    hipLaunchKernelGGL(k_gather_i, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, d_bl, c->d_pair_npp, bnpp);
    // staircase ends (bra-major [i][kseg]) and the estimated primitive quartets
    const long long nbk = (long long)nb * nkseg;
    int* jend = nullptr;
    int* jmid = nullptr;
    unsigned long long* d_red;
    HFK_TRY(t.alloc(&d_red, 2));
    HFK_HIP(hipMemsetAsync(d_red, 0, 2 * sizeof(unsigned long long), c->stream));
    HFK_TRY(plan_alloc(c, &jend, (size_t)std::max(1LL, nbk)));   // kept: the J'_cd layout reads it
    hipLaunchKernelGGL(k_staircase, dim3(grid_of(nbk, 256)), dim3(256), 0, c->stream, nb, nkseg, bq, kq, d_kfirst,
                       d_knpp, bnpp, e.kb == e.kk ? 1 : 0, ref ? 1 : 0, c->opt.screen_tau, jend, d_red);
    if (e.split) {
        double* smax;
        HFK_TRY(t.alloc(&smax, nk));
        HFK_TRY(t.alloc(&jmid, (size_t)std::max(1LL, nbk)));
        hipLaunchKernelGGL(k_suffix_max, dim3(grid_of(nkseg, 64)), dim3(64), 0, c->stream, nkseg, d_kfirst, ksch, smax);
        hipLaunchKernelGGL(k_split, dim3(grid_of(nbk, 256)), dim3(256), 0, c->stream, nb, nkseg, bsch, smax, d_kfirst,
                           jend, c->opt.fp32_threshold, jmid);
    }
    unsigned long long h_est = 0;
    HFK_HIP(hipMemcpyAsync(&h_est, d_red, sizeof(h_est), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    // Primitive ranges: a lane's share of a task is about per_lane primitive
    // quartets (at most 128), sized from the launch's total so that small
    // launches still fill the GPU (~1.3e5 lanes of work) and large ones keep
    // quartets whole.  Bra primitives are wave-uniform (LDS broadcast), so a
    // lane keeps as many of them as fit and the ket range is split.
    // Segments that need no split join the long multi-run tasks.
    const int L = e.la + e.lb + e.lc + e.ld;
    const long long floor_l = std::max(4, 32 >> L);   // cheap classes: amortise the digestion
    const int per_lane = (int)std::max(floor_l, std::min(128LL, (long long)h_est / 131072));
    const bool single = coop || hfk_class_grad_coop(e.la, e.lb, e.lc, e.ld);   // cooperative kernels read these
    auto whole = [&](int nbp, int nkp) { return !single && nbp <= HFK_BRA_STAGE && nbp * nkp <= per_lane; };
    std::vector<int> npiece((size_t)nbseg * nkseg, 1);
    for (int b = 0; b < nbseg; ++b)
        for (int s = 0; s < nkseg; ++s) {
            const int nbp = bnpp_seg[b], nkp = knpp[s];
            if (whole(nbp, nkp)) continue;
            int bstep = nbp, kstep = nkp;
            if (!coop && nbp * nkp > per_lane) {
                bstep = std::max(1, std::min(nbp, per_lane));
                kstep = std::max(1, std::min(nkp, per_lane / bstep));
            }
            if (!coop) bstep = std::min(bstep, HFK_BRA_STAGE);
            npiece[(size_t)b * nkseg + s] = ((nbp + bstep - 1) / bstep) * ((nkp + kstep - 1) / kstep);
        }
    int* d_npiece;
    HFK_TRY(plan_alloc(c, &d_npiece, npiece.size()));
    HFK_HIP(hipMemcpyAsync(d_npiece, npiece.data(), sizeof(int) * npiece.size(), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_merged, dim3(grid_of(nbk, 256)), dim3(256), 0, c->stream, nb, nkseg, jend, d_kfirst, d_knpp,
                       bnpp, single ? 1 : 0, per_lane, d_red + 1);
    // J'_cd entries: bra i and ket segment s own npiece x (jend - first)
    // entries from jbase[i][s], in bra-major order
    long long* jcdv;
    long long* jbase;
    HFK_TRY(t.alloc(&jcdv, (size_t)std::max(1LL, nbk)));
    HFK_TRY(plan_alloc(c, &jbase, (size_t)nbk + 1));
    hipLaunchKernelGGL(k_jcd_count, dim3(grid_of(nbk, 256)), dim3(256), 0, c->stream, nb, nkseg, nbseg, d_bfirst,
                       d_npiece, jend, d_kfirst, jcdv);
    HFK_TRY(scan_total(c, t, jcdv, jbase, nbk));
    unsigned long long h_merged = 0;
    long long h_njcd = 0;
    HFK_HIP(hipMemcpyAsync(&h_merged, d_red + 1, sizeof(h_merged), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipMemcpyAsync(&h_njcd, jbase + nbk, sizeof(h_njcd), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    e.njcd = h_njcd;
    // long tasks: about 40960 waves of work per launch, at least 256 and at
    // most 8192 kets each (each task stores its K' rows once, so longer
    // tasks store less; the costliest-first order keeps the tail short)
    const long long target = std::max(1, 40960 * 64 / block);
    const long long cmerge = std::max<long long>(std::max(chunk, 128),
                                                 std::min<long long>(8192, ((long long)h_merged + target - 1) / target));
    // tasks and runs of each precision part: count, offsets, write, order
    WalkArgs w;
    memset(&w, 0, sizeof w);
    w.nb = nb;
    w.nkseg = nkseg;
    w.split = e.split ? 1 : 0;
    w.single = single ? 1 : 0;
    w.coop = coop ? 1 : 0;
    w.per_lane = per_lane;
    w.chunk = chunk;
    w.cmerge = cmerge;
    w.bl = d_bl;
    w.bnpp = bnpp;
    w.kfirst = d_kfirst;
    w.knpp = d_knpp;
    w.soa_base = d_soa;
    w.jend = jend;
    w.jmid = jmid;
    w.jbase = jbase;
    const int nparts = e.split ? 2 : 1;
    long long *ntask[2], *nrun[2], *toff[2], *roff[2];
    long long T[2] = {0, 0}, R[2] = {0, 0};
    for (int part = 0; part < nparts; ++part) {
        HFK_TRY(t.alloc(&ntask[part], nb));
        HFK_TRY(t.alloc(&nrun[part], nb));
        HFK_TRY(t.alloc(&toff[part], (size_t)nb + 1));
        HFK_TRY(t.alloc(&roff[part], (size_t)nb + 1));
        w.part = part;
        w.ntask = ntask[part];
        w.nrun = nrun[part];
        hipLaunchKernelGGL(k_task_walk<false>, dim3(grid_of(nb, 128)), dim3(128), 0, c->stream, w);
        HFK_TRY(scan_total(c, t, ntask[part], toff[part], nb));
        HFK_TRY(scan_total(c, t, nrun[part], roff[part], nb));
        HFK_HIP(hipMemcpyAsync(&T[part], toff[part] + nb, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
        HFK_HIP(hipMemcpyAsync(&R[part], roff[part] + nb, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    }
    HFK_HIP(hipStreamSynchronize(c->stream));
    const long long ntot = T[0] + T[1], rtot = R[0] + R[1];
    // task and run indices are int; the K' reduction lists sort 2 entries per task
    if (2 * ntot > INT32_MAX || rtot > INT32_MAX) return hfk_set_error(HFK_EUNSUP, "task list too long");
    int4 *tasks_raw, *tasks;
    long long* cost;
    int* d_scal;   // [0] bra stage, [1] longest task
    HFK_TRY(t.alloc(&tasks_raw, (size_t)ntot));
    HFK_TRY(t.alloc(&cost, (size_t)ntot));
    HFK_TRY(t.alloc(&d_scal, 2));
    HFK_HIP(hipMemsetAsync(d_scal, 0, 2 * sizeof(int), c->stream));
    HFK_TRY(plan_alloc(c, &tasks, (size_t)ntot));
    KetRun* runs;
    HFK_TRY(plan_alloc(c, &runs, (size_t)rtot));
    w.tasks = tasks_raw;
    w.runs = runs;
    w.cost = cost;
    w.stage = d_scal;
    w.lenmax = d_scal + 1;
    {
        std::vector<long long> base = {0, 0, T[0], R[0]};   // (task, run) base of each part
        long long* d_base;
        HFK_TRY(to_dev(c, t, &d_base, base));
        long long tbase = 0;
        for (int part = 0; part < nparts; ++part) {
            long long *toff2, *roff2;
            HFK_TRY(t.alloc(&toff2, nb));
            HFK_TRY(t.alloc(&roff2, nb));
            if (nb > 0) {
                hipLaunchKernelGGL(k_add_offset, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, toff[part],
                                   d_base + 2 * part, toff2);
                hipLaunchKernelGGL(k_add_offset, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, roff[part],
                                   d_base + 2 * part + 1, roff2);
            }
            w.part = part;
            w.toff = toff2;
            w.roff = roff2;
            hipLaunchKernelGGL(k_task_walk<true>, dim3(grid_of(nb, 128)), dim3(128), 0, c->stream, w);
            // costliest first within the part
            if (T[part] > 0) {
                unsigned char *key, *key2;
                int *idx, *idx2;
                HFK_TRY(t.alloc(&key, (size_t)T[part]));
                HFK_TRY(t.alloc(&key2, (size_t)T[part]));
                HFK_TRY(t.alloc(&idx, (size_t)T[part]));
                HFK_TRY(t.alloc(&idx2, (size_t)T[part]));
                hipLaunchKernelGGL(k_cost_keys, dim3(grid_of(T[part], 256)), dim3(256), 0, c->stream, T[part],
                                   cost + tbase, key, idx);
                HFK_TRY(sort_pairs(c, t, key, key2, idx, idx2, T[part], 8));
                hipLaunchKernelGGL(k_gather_tasks, dim3(grid_of(T[part], 256)), dim3(256), 0, c->stream, T[part], idx2,
                                   tasks_raw + tbase, tasks + tbase);
            }
            tbase += T[part];
        }
    }
    HFK_HIP(hipGetLastError());
    int h_scal[2] = {0, 0};
    HFK_HIP(hipMemcpyAsync(h_scal, d_scal, sizeof(h_scal), hipMemcpyDeviceToHost, c->stream));
    HFK_HIP(hipStreamSynchronize(c->stream));
    e.bra_stage = h_scal[0];
    e.task_len_max = h_scal[1];
    e.ntask = ntot;
    e.ntask64 = T[0];
    e.d_tasks = tasks;
    e.d_runs = runs;
    // reduction lists: the tasks of each bra shell (K' rows, A then B per
    // task, in task order) and of each bra pair (J'_ab slots, in task order)
    const int nshell = (int)c->shells.size();
    int *pos_of, *kcnt, *bcnt, *koff, *boff;
    HFK_TRY(t.alloc(&pos_of, (size_t)c->npair));
    HFK_TRY(t.alloc(&kcnt, (size_t)nshell));
    HFK_TRY(t.alloc(&bcnt, (size_t)std::max(1, nb)));
    HFK_TRY(plan_alloc(c, &koff, (size_t)nshell + 1));
    HFK_TRY(plan_alloc(c, &boff, (size_t)nb + 1));
    HFK_HIP(hipMemsetAsync(kcnt, 0, sizeof(int) * nshell, c->stream));
    HFK_HIP(hipMemsetAsync(bcnt, 0, sizeof(int) * std::max(1, nb), c->stream));
    if (nb > 0) hipLaunchKernelGGL(k_pos_of, dim3(grid_of(nb, 256)), dim3(256), 0, c->stream, nb, d_bl, pos_of);
    unsigned *kkey, *kkey2, *bkey, *bkey2;
    int *kval, *kval2, *bval, *bval2;
    HFK_TRY(t.alloc(&kkey, (size_t)2 * ntot));
    HFK_TRY(t.alloc(&kkey2, (size_t)2 * ntot));
    HFK_TRY(t.alloc(&kval, (size_t)2 * ntot));
    HFK_TRY(t.alloc(&kval2, (size_t)2 * ntot));
    HFK_TRY(t.alloc(&bkey, (size_t)ntot));
    HFK_TRY(t.alloc(&bkey2, (size_t)ntot));
    HFK_TRY(t.alloc(&bval, (size_t)ntot));
    HFK_TRY(t.alloc(&bval2, (size_t)ntot));
    if (ntot > 0)
        hipLaunchKernelGGL(k_list_keys, dim3(grid_of(ntot, 256)), dim3(256), 0, c->stream, ntot, tasks, c->d_pair_A,
                           c->d_pair_B, pos_of, kkey, kval, bkey, bval, kcnt, bcnt);
    int kbits = 1, bbits = 1;
    while ((1LL << kbits) <= nshell) ++kbits;
    while ((1LL << bbits) <= nb) ++bbits;
    HFK_TRY(sort_pairs(c, t, kkey, kkey2, kval, kval2, 2 * ntot, kbits));
    HFK_TRY(sort_pairs(c, t, bkey, bkey2, bval, bval2, ntot, bbits));
    HFK_TRY(scan_total(c, t, kcnt, koff, nshell));
    HFK_TRY(scan_total(c, t, bcnt, boff, nb));
    int2* kcsr;
    int* jab_slot;
    HFK_TRY(plan_alloc(c, &kcsr, (size_t)2 * ntot));
    HFK_TRY(plan_alloc(c, &jab_slot, (size_t)ntot));
    if (ntot > 0)
        hipLaunchKernelGGL(k_list_out, dim3(grid_of(2 * ntot, 256)), dim3(256), 0, c->stream, ntot, NA, kval2, bval2,
                           kcsr, jab_slot);
    e.d_kcsr_off = koff;
    e.d_kcsr = kcsr;
    e.d_bcsr_off = boff;
    e.d_jab_slot = jab_slot;
    // K reduction jobs (HFK_KSEG entries of one shell component row) and rows
    {
        int *njob, *nrow, *joff, *rwoff;
        HFK_TRY(t.alloc(&njob, (size_t)nshell));
        HFK_TRY(t.alloc(&nrow, (size_t)nshell));
        HFK_TRY(t.alloc(&joff, (size_t)nshell + 1));
        HFK_TRY(t.alloc(&rwoff, (size_t)nshell + 1));
        hipLaunchKernelGGL(k_kjob_count, dim3(grid_of(nshell, 256)), dim3(256), 0, c->stream, nshell, koff, c->d_sh_bf,
                           njob, nrow);
        HFK_TRY(scan_total(c, t, njob, joff, nshell));
        HFK_TRY(scan_total(c, t, nrow, rwoff, nshell));
        int h_tot[2] = {0, 0};
        HFK_HIP(hipMemcpyAsync(&h_tot[0], joff + nshell, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HFK_HIP(hipMemcpyAsync(&h_tot[1], rwoff + nshell, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HFK_HIP(hipStreamSynchronize(c->stream));
        e.nkjob = h_tot[0];
        e.nkrow = h_tot[1];
        if (e.nkjob > 0) {
            HFK_TRY(plan_alloc(c, &e.d_kjob, (size_t)e.nkjob));
            HFK_TRY(plan_alloc(c, &e.d_krow, (size_t)e.nkrow));
            HFK_TRY(plan_alloc(c, &e.d_kseg, (size_t)e.nkjob * e.ncols));
            hipLaunchKernelGGL(k_kjob_write, dim3(grid_of(nshell, 256)), dim3(256), 0, c->stream, nshell, koff,
                               c->d_sh_bf, joff, rwoff, e.d_kjob, e.d_krow);
        }
    }
    // per ket position: descriptors, screen values, scales, position-major
    // record layout, ket segment
    {
        const int W = hfk_kdesc_words(NC, ND);
        std::vector<int> sh_atom(c->shells.size());
        for (size_t S = 0; S < c->shells.size(); ++S) sh_atom[S] = c->shells[S].atom;
        int* d_atom;
        HFK_TRY(to_dev(c, t, &d_atom, sh_atom));
        int *kdesc, *nonunit, *pos_len, *pos_npp, *pos_kseg;
        double *kposq, *kscale;
        long long* pos_soa;
        HFK_TRY(t.alloc(&nonunit, 1));
        HFK_HIP(hipMemsetAsync(nonunit, 0, sizeof(int), c->stream));
        HFK_TRY(t.alloc(&pos_soa, (size_t)nk));
        HFK_TRY(t.alloc(&pos_len, (size_t)nk));
        HFK_TRY(t.alloc(&pos_npp, (size_t)nk));
        HFK_TRY(t.alloc(&kscale, (size_t)(NC + ND) * nk));
        HFK_TRY(plan_alloc(c, &pos_kseg, (size_t)nk));
        if (!coop) {
            HFK_TRY(plan_alloc(c, &kdesc, (size_t)W * nk));
            HFK_TRY(plan_alloc(c, &kposq, (size_t)2 * nk));
        } else {
            HFK_TRY(t.alloc(&kdesc, (size_t)W * nk));
            HFK_TRY(t.alloc(&kposq, (size_t)2 * nk));
        }
        if (nk > 0)
            hipLaunchKernelGGL(k_ket_desc, dim3(grid_of(nk, 256)), dim3(256), 0, c->stream, nk, NC, ND, d_kl,
                               c->d_pair_A, c->d_pair_B, c->d_sh_bfp, c->d_sh_scale, d_atom, e.d_colmap,
                               c->d_pair_qhi, c->d_pair_sch, d_kfirst, nkseg, d_soa, d_knpp, kdesc, kposq, kscale,
                               nonunit, pos_soa, pos_len, pos_npp, pos_kseg);
        int h_nonunit = 0;
        HFK_HIP(hipMemcpyAsync(&h_nonunit, nonunit, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HFK_HIP(hipStreamSynchronize(c->stream));
        if (!coop) {
            e.d_kdesc = kdesc;
            e.d_kpos_q = kposq;
            if (h_nonunit) {
                HFK_TRY(plan_alloc(c, &e.d_kscale, (size_t)(NC + ND) * nk));
                HFK_HIP(hipMemcpyAsync(e.d_kscale, kscale, sizeof(double) * (NC + ND) * nk, hipMemcpyDeviceToDevice,
                                       c->stream));
            }
            // the class's position-major records (once per class)
            if (!c->d_ksoa[e.kk]) {
                HFK_TRY(dev_alloc(&c->d_ksoa[e.kk], (size_t)std::max(1LL, soa_base[nkseg])));
                hfk_launch_ket_soa(c->d_pp, c->d_qcomp, c->d_pair_pp0, d_kl, nk, pos_soa, pos_len, pos_npp, NF,
                                   nket_comp, c->d_ksoa[e.kk], c->stream);
                HFK_HIP(hipGetLastError());
            }
        }
        // the J'_cd reduction's layout: position -> segment, segment starts,
        // chunks of consecutive bras of one bra segment (enough chunks to
        // fill the GPU for small launches, at most ~128 plus one per segment
        // for large ones), tiles of ptile consecutive positions of one ket
        // segment
        const long long ch = std::max<long long>(16, ((long long)nb + 127) / 128);
        std::vector<int4> chunks;
        for (int b = 0; b < nbseg; ++b)
            for (long long i0 = bfirst[b]; i0 < bfirst[b + 1]; i0 += ch)
                chunks.push_back(make_int4(b, (int)i0, (int)std::min<long long>(bfirst[b + 1], i0 + ch), 0));
        const int ptile = std::max(1, 256 / nket_comp);
        std::vector<int2> tiles;
        for (int s = 0; s < nkseg; ++s)
            for (long long t0 = kfirst[s]; t0 < kfirst[s + 1]; t0 += ptile) tiles.push_back(make_int2(s, (int)t0));
        int* d_kf;
        int4* d_ch;
        int2* d_tl;
        HFK_TRY(plan_alloc(c, &d_kf, kfirst.size()));
        HFK_TRY(plan_alloc(c, &d_ch, chunks.size()));
        HFK_TRY(plan_alloc(c, &d_tl, tiles.size()));
        HFK_HIP(hipMemcpyAsync(d_kf, kfirst.data(), sizeof(int) * kfirst.size(), hipMemcpyHostToDevice, c->stream));
        if (!chunks.empty())
            HFK_HIP(hipMemcpyAsync(d_ch, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice, c->stream));
        if (!tiles.empty())
            HFK_HIP(hipMemcpyAsync(d_tl, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, c->stream));
        JcdLayout& J = e.jl;
        memset(&J, 0, sizeof J);
        J.tiles = d_tl;
        J.ntile = (int)tiles.size();
        J.ptile = ptile;
        J.pos_kseg = pos_kseg;
        J.kseg_first = d_kf;
        J.chunks = d_ch;
        J.jend = jend;
        J.jbase = jbase;
        J.npiece = d_npiece;
        J.nket = nk;
        J.nkseg = nkseg;
        J.nchunk = (int)chunks.size();
        J.nket_comp = nket_comp;
        J.same = e.kb == e.kk;
    }
    HFK_HIP(hipStreamSynchronize(c->stream));
    return HFK_OK;
}
