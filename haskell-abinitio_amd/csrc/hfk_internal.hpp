// Internal interfaces between the host C-ABI layer and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfk_device.hpp"

#define HFK_MAXCOMP 6        // Cartesian components of a d shell
#define HFK_LMAX 2           // highest shell angular momentum with generated kernels

// Everything a quartet kernel needs.  Pair p: shells pair_A[p] (the one
// with the larger angular momentum) and pair_B[p]; its primitive pairs are
// pp[(pair_pp0[p] + i) * PP_STRIDE], i < pair_npp[p].
struct QuartetArgs {
    const double* pp;
    const double* qcomp;
    const int* pair_pp0;
    const int* pair_npp;
    const int* pair_A;
    const int* pair_B;
    const double* pair_qhi;     // max reference screen value over the pair (NaN -> +inf)
    const double* pair_sch;     // Schwarz bound sqrt(max |(ab|ab)|), for the precision split
    const int* bra_list;        // pair indices of the bra class
    const int* ket_list;        // pair indices of the ket class
    long long nbra, nket;
    int same;                   // bra class == ket class: only j <= i
    long long nquart;
    const int* sh_bf;           // [nshell][HFK_MAXCOMP] basis-function index or -1
    const int* sh_bfp;          // same with -1 replaced by n (the zero row of padded D/J/K)
    const double* sh_scale;     // [nshell][HFK_MAXCOMP] component coefficient scale
    BoysTables bt;
    double qtau;                // reference primitive screen threshold (0 = off)
    double sch_cut;             // skip quartets with Schwarz bound below (0 = off)
    double fp32_thr;            // precision split threshold on the Schwarz bound
    int prec_sel;               // 0: all quartets; 1: bound >= fp32_thr; 2: bound < fp32_thr
    int n;                      // number of basis functions
    const double* D;            // [n+1][n+1] density (symmetric), row/column n zero
    double* J;                  // [n+1][n+1] accumulator J' (symmetrised on output)
    double* Jk;                 // ket-side sums J'_cd by ket position: [jk_copies][nket][NKET]
    int jk_copies;              // (workgroup w adds into copy w % jk_copies; scattered into J)
    double* K;                  // [n+1][n+1] accumulator K'
    double* flat;               // flat ERI output (calcIntegrals order)
    const long long* flat_base; // [n][n] rank of (i,j,i,j)
    double* diag;               // per-pair max |(ab|ab)| output (Schwarz setup)
    // task list of the tiled kernels: {bra pair, first run, end run, primitive
    // ranges (b0 << 24 | b1 << 16 | k0 << 8 | k1; k1 = 255: all of the ket's)}
    const int4* tasks;
    const int2* runs;           // [begin, end) positions in ket_idx, walked in order
    const int* ket_idx;         // ket pair indices (the ket class list)
    long long ntask;
    int lds_rows;               // K rows staged in LDS (0: global atomics)
    int block;                  // workgroup size of the tiled kernel (64 or 256)
    int bra_stage;              // longest bra primitive-pair range of the tasks (LDS staging)
    int task_len_max;           // longest ket run of the tasks
    // nuclear gradient
    const int* sh_atom;         // [nshell] atom of each shell
    double* grad;               // [natom][3] accumulator
    int natom;
    int grad_lds;               // per-workgroup atom accumulator in LDS (0: global atomics)
};

enum { HFK_OP_JK = 0, HFK_OP_FLAT = 1, HFK_OP_DIAG = 2, HFK_OP_GRAD = 3 };

// Classes with a d shell and more than 54 contracted integrals per quartet
// run J/K on the cooperative wave kernel (64-lane
// workgroups, FP64, whole quartets per task); gen_eri.py uses the same bound.
inline bool hfk_class_coop(int la, int lb, int lc, int ld)
{
    const int nv = (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2) * ((ld + 1) * (ld + 2) / 2);
    return (la >= 2 || lb >= 2 || lc >= 2 || ld >= 2) && nv > 54;
}

// Classes whose gradient runs on the cooperative kernel (d classes with more
// than 18 integrals per quartet, and any class with more than 54); gen_eri.py
// uses the same bounds.  That kernel needs tasks of one run with an explicit
// ket primitive range (uniform over the workgroup).
inline bool hfk_class_grad_coop(int la, int lb, int lc, int ld)
{
    const int nv = (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2) * ((ld + 1) * (ld + 2) / 2);
    return ((la >= 2 || lb >= 2 || lc >= 2 || ld >= 2) && nv > 18) || nv > 54;
}

// The per-lane J/K kernel stages the task's bra primitive-pair records and
// their per-component screen values in LDS: at most HFK_BRA_STAGE primitive
// pairs per task (the planner splits longer bra ranges).
#define HFK_BRA_STAGE 64
inline size_t hfk_bra_stage_bytes(int la, int lb, int nprim)
{
    return sizeof(double) * nprim * (16 + (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2));
}

// Launch one class kernel.  classes are (la,lb,lc,ld) with la>=lb, lc>=ld,
// (la,lb) >= (lc,ld).  prec: 0 = FP64, 1 = FP32.  boys: 0 reference, 1 accurate.
// mode (J/K only): 0 main pass, 1 screen-correction pass.
int hfk_launch_quartets(int la, int lb, int lc, int ld, int op, int prec, int boys, int mode,
                        const QuartetArgs& a, hipStream_t s);

// Two-electron gradient contribution of one class over a plan entry's task list
// (FP64), accumulated into a.grad.
int hfk_launch_grad(int la, int lb, int lc, int ld, int boys, int mode, const QuartetArgs& a, hipStream_t s);

// Setup kernels (hfk_setup.hip)
struct ShellDev {
    const double* center;   // [nshell][3]
    const int* L;
    const int* prim0;       // first primitive in expo[]/coef[]
    const int* nprim;
    const double* expo;
    const double* coef;     // base contraction coefficients
    const double* scale;    // [nshell][HFK_MAXCOMP]
};

void hfk_launch_pair_setup(const ShellDev& sh, int npair, const int* pair_A, const int* pair_B,
                           const int* pair_pp0, const int* pair_q0, double* pp, double* qcomp,
                           double* pair_qhi, hipStream_t s);
void hfk_launch_boys_tables(const double* glx, const double* glw, double* ref64, float* ref32,
                            double* acc64, float* acc32, hipStream_t s);
void hfk_launch_fill(double* p, double v, long long n, hipStream_t s);
// Exact counts over a launch's task list (surviving quartets only): shell,
// primitive and canonical function quartets, out[0..2] (atomic adds).  Shell and
// function quartets are counted once per quartet, primitive quartets per
// primitive range.
void hfk_launch_count_tasks(const QuartetArgs& a, int na, int nb, int nc, int nd, unsigned long long* out,
                            hipStream_t s);
// J, K padded (leading dimension n+1)
void hfk_launch_symmetrize_g(int n, const double* J, const double* K, const double* H, double* G_flat,
                             double* F_full, hipStream_t s);
// One J/K launch's ket-side sums in Jk: [copies][nket][nc * nd] from offset off,
// for the ket pairs ket_list[0..nket)
struct JkScatter {
    const int* ket_list;
    long long nket;
    long long off;
    int nc, nd;
};
// J'[c][d] += sum over copies of Jk[off + (copy * nket + pos) * nc * nd + z] for
// every launch e of ent[0..nent) (one kernel; ld = n + 1)
void hfk_launch_scatter_jk(double* J, const double* Jk, int copies, const JkScatter* ent, int nent, long long max_count,
                           const int* pair_A, const int* pair_B, const int* sh_bfp, int ld, hipStream_t s);
// flat -> padded (n+1)x(n+1) full, row/column n zero
void hfk_launch_flat_to_padded(int n, const double* flat, double* full, hipStream_t s);
// padded D from the flat density; J', K' ((n+1)^2) and Jk (jk_len) zeroed; one launch
void hfk_launch_fock_prepare(int n, const double* D_flat, double* D, double* J, double* K, double* Jk, long long jk_len,
                             hipStream_t s);
void hfk_launch_flat_to_full(int n, const double* flat, double* full, hipStream_t s);
void hfk_launch_full_to_flat(int n, const double* full, double* flat, hipStream_t s);

// One-electron integrals (hfk_onee.hip)
struct OneEArgs {
    int n;
    const double* center;   // [n][3]
    const int* lxyz;        // [n][3]
    const int* prim_off;    // [n+1]
    const double* coef;
    const double* expo;
    int natom;
    const double* Z;
    const double* R;        // [natom][3]
    BoysTables bt;
    int boys;
    const int* vpair;       // flat pair indices grouped by l_i + l_j = 0..4
    const int* vpair_off;   // [6] group offsets into vpair (host pointer)
};
void hfk_launch_overlap(const OneEArgs& a, double* S_flat, hipStream_t s);
void hfk_launch_hcore(const OneEArgs& a, double* H_flat, hipStream_t s);
// One-electron gradient: sum D_ij dH_ij/dR - W_ij dS_ij/dR (flat D, W) into grad[natom][3].
void hfk_launch_grad_onee(const OneEArgs& a, const int* atom_of_bf, const double* D_flat, const double* W_flat,
                          double* grad, hipStream_t s);

// Dense linear algebra (hfk_linalg.hip): row-major FP64, MFMA.
// C = alpha * op(A) * op(B) + beta * C, with op = transpose when ta/tb.
void hfk_launch_dgemm(int m, int n, int k, double alpha, const double* A, int lda, int ta,
                      const double* B, int ldb, int tb, double beta, double* C, int ldc, hipStream_t s);
void hfk_launch_scale_cols(int n, double* U, const double* w, int pow_mode, hipStream_t s);
void hfk_launch_reduce(long long n, const double* x, const double* y, int mode, double* part, double* out, hipStream_t s);
void hfk_launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s);
void hfk_launch_antisym(int n, const double* T, double* E, hipStream_t s);
void hfk_launch_guess_density(int n, int occ, double* Dflat, hipStream_t s);
