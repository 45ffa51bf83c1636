// Host-side context and helpers shared by the hfk host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "hfk.h"
#include "hfk_internal.hpp"

#define HFK_NPC ((HFK_LMAX + 1) * (HFK_LMAX + 2) / 2)   // pair classes (la >= lb)

int hfk_set_error(int code, const char* fmt, ...);

#define HFK_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return hfk_set_error(HFK_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                 __LINE__);                                                             \
    } while (0)

#define HFK_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != HFK_OK) return r_; \
    } while (0)

struct HostShell {
    int atom;
    double R[3];
    int L;
    int prim0, nprim;           // into sh_expo / sh_coef
    int bf[HFK_MAXCOMP];        // basis function of each canonical component, or -1
    double scale[HFK_MAXCOMP];  // coefficient of the component / base coefficient
};

// Segment of a pair-class list: positions [first, end), npp primitive pairs each.
struct ClassSeg {
    long long first, end;
    int npp;
};

struct PlanEntry {
    int kb, kk;                 // bra / ket pair class
    int la, lb, lc, ld;
    int split;                  // tasks [0, ntask64) run in FP64, [ntask64, ntask) in FP32
    long long nquart;           // launched quartet slots
    long long shell_q[2], prim_q[2], func_q[2];   // counted survivors of the FP64 / FP32 part
    double ms[2];               // last timed duration of each part's launch
    double grad_ms;             // its launch in the last gradient (all tasks, FP64)
    int mode;                   // 0 main pass, 1 screen-correction pass (shares the main tasks)
    int4* d_tasks;              // tasks {bra pair, first run, end run, primitive ranges}
    KetRun* d_runs;             // runs [begin, end) of positions in d_kets, with their J'_cd entries
    int* d_kets;                // ket pair indices the tasks refer to (the ket class list)
    long long ntask, ntask64;
    int bra_stage;              // longest bra primitive-pair range of the tasks
    int task_len_max;           // longest ket sequence of the tasks
    int solo;                   // its last timed run had the GPU to itself
    // J/K slots (written by the launch with plain stores) and their layout
    int ncols;                  // compact K columns: functions of the ket class's shells
    int* d_colmap;              // [n+1] function -> column (n -> ncols)
    int* d_collist;             // [ncols]
    // per-lane kernels: ket data by class-list position (QuartetArgs)
    int* d_kdesc;               // [hfk_kdesc_words][nket]
    double* d_kpos_q;           // [2][nket]
    double* d_kscale;           // [NC + ND][nket], null when every scale is 1
    int drows;                  // the J/K kernel stages the bra shells' density rows
    double* d_kslot;            // [ntask][NA+NB][ncols]
    double* d_jab;              // [ntask][NA*NB]
    double* d_jcd;              // [njcd][NC*ND]
    long long njcd;
    JcdLayout jl;               // device pointers of the J'_cd layout
    JKPart jkp;                 // its partial J', K' (reduced on its stream after the launch)
    int* d_kcsr_off;            // [nshell+1]
    int2* d_kcsr;               // {task, row offset}
    int4* d_kjob;               // K reduction jobs (JKPart.kjob)
    int4* d_krow;               // K reduction rows (JKPart.krow)
    double* d_kseg;             // [nkjob][ncols] first-pass sums
    int nkjob, nkrow;
    int* d_bcsr_off;            // [nbra+1]: J'_ab slots of each bra position
    int* d_jab_slot;            // [ntask]
    size_t slot_bytes;          // device memory of the slots
    std::vector<void*> owned;   // device allocations of this entry
};

struct hfk_ctx {
    std::mutex mu;
    int device = -1;            // HIP device current at hfk_create
    int n = 0, natom = 0;
    hfk_options opt;
    // basis as given
    std::vector<double> center, coef, expo, Z, R;
    std::vector<int> lxyz, prim_off, atom_of_bf;
    // shells
    std::vector<HostShell> shells;
    std::vector<double> sh_expo, sh_coef;
    // pairs
    int npair = 0;
    std::vector<int> h_pair_A, h_pair_B, h_pair_npp;
    // pair classes (built on the device, hfk_plan.hip): their lengths and
    // their segments (runs of one primitive-pair count)
    long long class_count[HFK_NPC] = {0};
    std::vector<ClassSeg> class_segs[HFK_NPC];
    int* d_class[HFK_NPC] = {nullptr};
    // position-major copies of the classes' primitive-pair records, for the
    // classes that are kets of per-lane launches (KetRun.soa); built with the plan
    double* d_ksoa[HFK_NPC] = {nullptr};
    // device data
    hipStream_t stream = nullptr;
    // J/K launches of one Fock build run concurrently on these (forked from
    // and joined into `stream`); default priority
    static constexpr int NAUX = 3;
    hipStream_t aux[NAUX] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[NAUX] = {nullptr, nullptr, nullptr};
    int concurrent = 1;
    // the last converged hfk_scf: the density its last Fock build used and that
    // (raw, not extrapolated) Fock matrix, flat; hfk_gradient builds F(D) from them
    double* d_scf_D = nullptr;
    double* d_scf_F = nullptr;
    bool scf_cached = false;
    int first_launch = -1;      // plan entry that runs alone first (-1: not yet measured)
    int serial_runs = 0;        // serial J/K runs since planning (the first is cold)
    bool timed_serial = false;  // the last timed run launched serially
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* blas = nullptr;                 // rocblas_handle
    double* h_pin = nullptr;              // pinned host: HFK_PIN_DOUBLES scalars, then one int status
#define HFK_PIN_DOUBLES 64
    std::vector<void*> allocs;
    double *d_glx = nullptr, *d_glw = nullptr;
    double *d_boys_ref64 = nullptr, *d_boys_acc64 = nullptr;
    float *d_boys_ref32 = nullptr, *d_boys_acc32 = nullptr;
    BoysTables bt;
    double *d_center = nullptr, *d_coef = nullptr, *d_expo = nullptr, *d_Z = nullptr, *d_R = nullptr;
    int *d_lxyz = nullptr, *d_prim_off = nullptr;
    double *d_sh_center = nullptr, *d_sh_scale = nullptr, *d_sh_expo = nullptr, *d_sh_coef = nullptr;
    int *d_sh_L = nullptr, *d_sh_prim0 = nullptr, *d_sh_nprim = nullptr, *d_sh_bf = nullptr, *d_sh_bfp = nullptr;
    int *d_pair_A = nullptr, *d_pair_B = nullptr, *d_pair_pp0 = nullptr, *d_pair_npp = nullptr;
    double *d_pp = nullptr, *d_qcomp = nullptr, *d_pair_qhi = nullptr, *d_pair_sch = nullptr, *d_pair_diag = nullptr;
    double* d_D = nullptr;      // padded density of the Fock build
    // per-launch partial J', K' of the J/K slots: [plan entries][n+1][n+1] each
    double* d_kpart = nullptr;
    double* d_jpart = nullptr;
    size_t slot_bytes = 0;          // device memory of all J/K slots
    double *d_flatA = nullptr, *d_flatB = nullptr, *d_flatC = nullptr;
    int *d_sh_atom = nullptr, *d_atom_of_bf = nullptr;   // gradient (uploaded on first use)
    // device arrays of the current plan, bump-allocated from chunks (plan_alloc);
    // freed together when the plan is rebuilt or the context destroyed
    std::vector<void*> plan_chunks;
    char* plan_cur = nullptr;
    size_t plan_left = 0;
    double* d_gwork = nullptr;   // gradient workspace: 3 n*n + 2 n(n+1)/2 + 3 natom doubles (first use)
    int* d_vpair = nullptr;                 // function pairs grouped by l_i + l_j (one-electron V)
    std::vector<int> vpair_off;             // [6]
    hfk_stats stats = {};
    std::vector<PlanEntry> plan;
    std::vector<hipEvent_t> events;
    int timed_op = 0;
    bool counted = false;       // plan[].shell_q/prim_q/func_q and stats.func_quartets are current
    int kets_per_lane = 1;      // kets per lane in a tiled J/K task
};

// Workgroup size of a J/K launch: 256 lanes (four waves sharing one set of
// staged K rows) when the rows are long enough to limit occupancy through LDS
// and the class runs per lane; otherwise 64.
inline int hfk_plan_block(const hfk_ctx* c, const PlanEntry& e)
{
    return (c->n >= 200 && !hfk_class_coop(e.la, e.lb, e.lc, e.ld)) ? 256 : 64;
}
// LDS a per-lane workgroup of `block` lanes may take so that four waves per
// SIMD (16 per CU) fit in a CU's 160 KB; at most 64 KB.
inline size_t hfk_lds_budget(int block)
{
    const size_t b = (size_t)160 * 1024 * block / (16 * 64);
    return b < 64 * 1024 ? b : 64 * 1024;
}

template <typename T>
int dev_alloc(T** p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * n);
    if (e != hipSuccess) return hfk_set_error(HFK_ENOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * n, hipGetErrorString(e));
    return HFK_OK;
}

template <typename T>
int dev_alloc_tracked(hfk_ctx* c, T** p, size_t n)
{
    int rc = dev_alloc(p, n);
    if (rc == HFK_OK) c->allocs.push_back((void*)*p);
    return rc;
}

template <typename T>
int dev_upload(hfk_ctx* c, T** p, const std::vector<T>& v)
{
    int rc = dev_alloc(p, v.size());
    if (rc != HFK_OK) return rc;
    c->allocs.push_back((void*)*p);
    if (!v.empty()) {
        hipError_t e = hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hfk_set_error(HFK_EHIP, "upload: %s", hipGetErrorString(e));
    }
    return HFK_OK;
}

template <typename T>
void dev_free(hfk_ctx* c, T* p)
{
    if (!p) return;
    hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->allocs.size(); ++i)
        if (c->allocs[i] == (void*)p) {
            c->allocs.erase(c->allocs.begin() + i);
            break;
        }
    hipFree((void*)p);
}

// dev_alloc used for context-owned buffers is tracked too
#define dev_alloc(p, n) dev_alloc_tracked(c, p, n)

// A device array owned by the current plan: bump-allocated (256-byte aligned)
// from chunks of at least 16 MB, so planning a small molecule makes a few
// allocations instead of one per array; hfk_plan_free_arrays releases them.
template <typename T>
int plan_alloc(hfk_ctx* c, T** p, size_t n)
{
    constexpr size_t ALIGN = 256, CHUNK = 16u << 20;
    const size_t bytes = ((std::max<size_t>(n, 1) * sizeof(T) + ALIGN - 1) / ALIGN) * ALIGN;
    if (bytes > c->plan_left) {
        const size_t sz = std::max(bytes, CHUNK);
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sz);
        if (e != hipSuccess) return hfk_set_error(HFK_ENOMEM, "hipMalloc(%zu bytes): %s", sz, hipGetErrorString(e));
        c->plan_chunks.push_back(q);
        c->plan_cur = (char*)q;
        c->plan_left = sz;
    }
    *p = (T*)c->plan_cur;
    c->plan_cur += bytes;
    c->plan_left -= bytes;
    return HFK_OK;
}
void hfk_plan_free_arrays(hfk_ctx* c);

int hfk_ctx_init_device(hfk_ctx* c);
// hfk_noalloc.hip: this device's copy of the fixed scratch array (once per device)
int hfk_noalloc_setup(int device);
void hfk_ctx_free_device(hfk_ctx* c);
int hfk_ctx_classify_pairs(hfk_ctx* c);
void hfk_pair_class_l(int k, int* la, int* lb);
QuartetArgs hfk_ctx_args(hfk_ctx* c);
int hfk_ctx_run_quartets(hfk_ctx* c, QuartetArgs a, int op);
int hfk_ctx_plan(hfk_ctx* c);
// hfk_plan.hip: the task and reduction lists of one plan entry, built on the
// device (colmap / ncols set before); sqrt of n doubles on the device
int hfk_plan_entry_device(hfk_ctx* c, PlanEntry& e);
void hfk_launch_sqrt(int n, const double* in, double* out, hipStream_t s);
void hfk_plan_args(hfk_ctx* c, const PlanEntry& e, QuartetArgs& a);
int hfk_ctx_collect_times(hfk_ctx* c);
// calcIntegrals into a host array of out_len = M(M+1)/2 (caller holds the lock)
int hfk_ctx_eri_flat(hfk_ctx* c, double* out, size_t out_len);
// extra_cut > 0 also skips shell quartets with Schwarz product below it
int hfk_ctx_fock_device(hfk_ctx* c, const double* H_dev, const double* D_dev_flat, double* F_dev_flat,
                        double extra_cut = 0.0);
int hfk_ctx_overlap_dev(hfk_ctx* c, double* S_dev);
int hfk_ctx_hcore_dev(hfk_ctx* c, double* H_dev);
OneEArgs hfk_onee_args(hfk_ctx* c);

// linear algebra helpers (hfk_linalg_host.hip)
int hfk_linalg_init(hfk_ctx* c);
void hfk_linalg_destroy(hfk_ctx* c);
// symmetric eigen-solve on device: A (n x n row-major, overwritten by
// eigenvectors as ROWS: A[k][:] = k-th eigenvector), w ascending.
// E (n) and info (2) are device workspace; info_host (pinned) or NULL: see hfk_scf.hip
int hfk_dev_syev(hfk_ctx* c, int n, double* A, double* w, double* E, int* info, int* info_host);
