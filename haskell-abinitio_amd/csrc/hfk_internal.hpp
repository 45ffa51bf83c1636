// Internal interfaces between the host C-ABI layer and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfk_device.hpp"

#define HFK_MAXCOMP 6        // Cartesian components of a d shell
#define HFK_LMAX 2           // highest shell angular momentum with generated kernels
#define HFK_KSEG 32          // K-slot rows summed per first-pass job of the K reduction

// Ket run of a task: positions [x, y) of the ket class list, walked in
// order; jcd is the J'_cd slot entry of position x (position p writes entry
// jcd + p - x, NKET doubles each).  The positions of a run lie in one ket
// segment (equal primitive-pair count npp, len positions); field f of
// primitive-pair ik of position p is ksoa[soa + p + (ik * NF + f) * len]
// in the class's position-major copy of the records (per-lane kernels).
struct KetRun {
    int x, y;
    long long jcd;
    long long soa;
    int len, npp;
};

// Fields of a position-major ket record: the PP_STRIDE fields of the
// primitive-pair record, then its per-component screen values.
constexpr int hfk_ket_fields(int lc, int ld) { return 16 + (lc + 1) * (lc + 2) / 2 * ((ld + 1) * (ld + 2) / 2); }
// Per-position ket descriptor of a launch, kdesc[x * nket + p] for word x:
// the padded basis-function indices of shells C (NC words) and D (ND), their
// compact K columns (NC, ND), a flag word (bit 0: C == D), the atoms of C
// and D.
constexpr int hfk_kdesc_words(int nc, int nd) { return 2 * (nc + nd) + 3; }

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
    const int* sh_bfp;          // same with -1 replaced by n (the zero row of padded D)
    const double* sh_scale;     // [nshell][HFK_MAXCOMP] component coefficient scale
    BoysTables bt;
    double qtau;                // reference primitive screen threshold (0 = off)
    double sch_cut;             // skip quartets with Schwarz bound below (0 = off)
    double fp32_thr;            // precision split threshold on the Schwarz bound
    int prec_sel;               // 0: all quartets; 1: bound >= fp32_thr; 2: bound < fp32_thr
    int n;                      // number of basis functions
    const double* D;            // [n+1][n+1] density (symmetric), row/column n zero
    // J/K digestion: every task writes its partial results to its own slots
    // with plain stores (no atomics); fixed-order reductions sum them later
    double* kslot;              // [task][(NA+NB) rows][ncols]: K' rows of the bra shells
    double* jab;                // [slot][NBRA]: J'_ab of the bra pair; task t writes slot jab_slot[t]
    const int* jab_slot;        // [task]: the tasks of one bra pair have consecutive slots
    double* jcd;                // J'_cd entries (KetRun.jcd), [entry][NKET]
    const int* colmap;          // [n+1] basis function -> compact K column; n -> ncols (discarded)
    int ncols;                  // K columns of the launch (functions of its ket shells)
    long long task0;            // index of tasks[0] in the launch's task list (slot numbering)
    double* flat;               // flat ERI output (calcIntegrals order)
    const long long* flat_base; // [n][n] rank of (i,j,i,j)
    double* diag;               // per-pair max |(ab|ab)| output (Schwarz setup)
    // task list of the tiled kernels: {bra pair, first run, end run, primitive
    // ranges (b0 << 24 | b1 << 16 | k0 << 8 | k1; k1 = 255: all of the ket's)}
    const int4* tasks;
    const KetRun* runs;         // runs of the tasks
    const int* ket_idx;         // ket pair indices (the ket class list)
    const double* ksoa;         // position-major ket records of the ket class (KetRun.soa)
    const int* kdesc;           // [hfk_kdesc_words][nket] per-position ket descriptor
    const double* kpos_q;       // [2][nket]: pair_qhi, pair_sch of each ket position
    const double* kscale;       // [NC + ND][nket] component scales of C, D (null: all 1)
    const int* collist;         // [ncols] basis function of each compact column
    int drows;                  // stage the bra shells' density rows in LDS (compact columns)
    long long ntask;
    int block;                  // workgroup size of the tiled kernel (64 or 256)
    int bra_stage;              // longest bra primitive-pair range of the tasks (LDS staging)
    int task_len_max;           // longest ket sequence of the tasks
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
// than 18 integrals per quartet; (pp|pp) splits its bra components over
// per-lane launches instead); gen_eri.py uses the same bounds.  That kernel needs tasks of one run with an explicit
// ket primitive range (uniform over the workgroup).
inline bool hfk_class_grad_coop(int la, int lb, int lc, int ld)
{
    const int nv = (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2) * ((ld + 1) * (ld + 2) / 2);
    return (la >= 2 || lb >= 2 || lc >= 2 || ld >= 2) && nv > 18;
}

// The per-lane J/K kernel stages the task's bra primitive-pair records and
// their per-component screen values in LDS: at most HFK_BRA_STAGE primitive
// pairs per task (the planner splits longer bra ranges).
#define HFK_BRA_STAGE 64
inline size_t hfk_bra_stage_bytes(int la, int lb, int nprim)
{
    return sizeof(double) * nprim * (16 + (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2));
}

// Runs of one J/K task at most (the per-lane kernel stages them in LDS:
// J'_cd entry, record base, first position, start in the task's sequence,
// segment length, primitive-pair count).
#define HFK_TASK_RUNS 64
inline size_t hfk_run_stage_bytes() { return (2 * sizeof(long long) + 4 * sizeof(int)) * (HFK_TASK_RUNS + 1); }
// LDS bytes of a per-lane J/K workgroup: K' rows of the bra shells over the
// launch's ncols compact columns (+ one discarded), the cross-wave J'_ab
// sums, the task's runs (first ket, entry, start), the staged bra
// primitive pairs.
inline size_t hfk_jk_lds_bytes(int la, int lb, int ncols, int bra_stage, int drows)
{
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    return sizeof(double) * ((size_t)(drows ? 2 : 1) * (na + nb) * (ncols + 1) + 4 * (size_t)na * nb) +
           hfk_run_stage_bytes() + hfk_bra_stage_bytes(la, lb, bra_stage);
}
// LDS bytes of a per-lane gradient workgroup: the per-atom accumulator, the
// bra shells' density rows (drows), the task's runs, the staged bra
// primitive pairs.
inline size_t hfk_grad_lds_bytes(int la, int lb, int ncols, int natom, int drows, int bra_stage)
{
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    return sizeof(double) * (3 * (size_t)natom + (drows ? (size_t)(na + nb) * (ncols + 1) : 0)) + hfk_run_stage_bytes() +
           hfk_bra_stage_bytes(la, lb, bra_stage);
}
// Dynamic LDS of a cooperative J/K workgroup (its static tables come on top).
inline size_t hfk_coop_jk_lds_bytes(int la, int lb, int ncols)
{
    return sizeof(double) * ((size_t)((la + 1) * (la + 2) / 2 + (lb + 1) * (lb + 2) / 2) * (ncols + 1) + 64);
}

// Launch one class kernel.  classes are (la,lb,lc,ld) with la>=lb, lc>=ld,
// (la,lb) >= (lc,ld).  prec: 0 = FP64, 1 = FP32.  boys: 0 reference, 1 accurate.
// mode (J/K only): 0 main pass, 1 screen-correction pass.
int hfk_launch_quartets(int la, int lb, int lc, int ld, int op, int prec, int boys, int mode,
                        const QuartetArgs& a, hipStream_t s);

// Two-electron gradient contribution of one class over the tasks in a,
// accumulated into a.grad.  prec: 0 FP64, 1 FP32 (the per-lane kernels; the
// cooperative d-class kernel is FP64 and ignores it).
int hfk_launch_grad(int la, int lb, int lc, int ld, int boys, int mode, int prec, const QuartetArgs& a, hipStream_t s);

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

// Each pair's records sorted by descending QHI (stable) into out, their
// per-component screen values into qout in the same order.
void hfk_launch_pair_sort(int npair, const int* pair_pp0, const int* pair_npp, const int* pair_q0, long long nq_tot,
                          const double* pp, const double* qcomp, double* out, double* qout, hipStream_t s);
void hfk_launch_pair_setup(const ShellDev& sh, int npair, const int* pair_A, const int* pair_B,
                           const int* pair_pp0, const int* pair_q0, double* pp, double* qcomp,
                           double* pair_qhi, hipStream_t s);
// Position-major copy of a ket class's primitive-pair records: for position
// j (pair klist[j], pos_npp[j] primitive pairs, segment length pos_len[j]),
// field f < PP_STRIDE of primitive ik goes to out[pos_soa[j] + j + (ik * nf +
// f) * pos_len[j]], its ncomp screen values to the fields after.
void hfk_launch_ket_soa(const double* pp, const double* qcomp, const int* pair_pp0, const int* klist, int nket,
                        const long long* pos_soa, const int* pos_len, const int* pos_npp, int nf, int ncomp,
                        double* out, hipStream_t s);
void hfk_launch_boys_tables(const double* glx, const double* glw, double* ref64, float* ref32,
                            double* acc64, float* acc32, hipStream_t s);
void hfk_launch_fill(double* p, double v, long long n, hipStream_t s);
// Exact counts over a launch's task list (surviving quartets only): shell,
// primitive and canonical function quartets, out[0..2] (atomic adds).  Shell and
// function quartets are counted once per quartet, primitive quartets per
// primitive range.
void hfk_launch_count_tasks(const QuartetArgs& a, int na, int nb, int nc, int nd, unsigned long long* out,
                            hipStream_t s);
// Fixed-order reductions of the J/K slots (hfk_setup.hip).
// J'_cd of one launch: for each chunk of consecutive bras of one bra
// segment and each ket position, the sum over the chunk's bras that cover
// the position (bra-list order) and over the primitive pieces, into
// part[chunk][pos][k] (plain stores); the J' partial then sums the chunks in
// order.
struct JcdLayout {
    const double* jcd;          // the launch's entries
    double* part;               // [nchunk][nket][nket_comp]
    const int* pos_kseg;        // [nket] ket segment of each position
    const int* kseg_first;      // [nkseg + 1]
    const int4* chunks;         // [nchunk] {bra segment, first bra, end bra, -}
    const int2* tiles;          // [ntile] {ket segment, first position}: ptile positions of one segment
    int ntile, ptile;
    const int* jend;            // [nbra][nkseg] exclusive end position of bra i's kets in the segment
    const long long* jbase;     // [nbra][nkseg] entry of the first piece
    const int* npiece;          // [nbseg][nkseg] primitive pieces per (bra, ket) quartet
    int nket, nkseg, nchunk, nket_comp, same;
};
void hfk_launch_reduce_jcd(const JcdLayout& L, hipStream_t s);
// The per-launch partial J' and K' ((n+1) x (n+1) each; every build writes
// the same elements, the others stay zero from planning):
//   K: for each basis function r of a shell S, the sum over the tasks whose
//      bra holds S (task order) of their slot rows, scattered to the launch's
//      columns;
//   J: the bra side (J'_ab slots of each bra pair, in order) and the ket
//      side (the J'_cd chunk sums, in chunk order) at the pairs' function
//      elements; for a class
//      paired with itself both sides of a pair go to one element in one sum.
// The Fock build then sums the launches' partials in plan order.
struct JKPart {
    const double* kslot;        // [task][nrow][ncols]
    const int* csr_off;         // [nshell + 1] into csr
    const int2* csr;            // {task, row offset (0: shell A of the bra, NA: shell B)}
    const int* collist;         // [ncols] basis function of each compact column
    int ncols, nrow, nshell;
    // K rows in two passes: job = HFK_KSEG consecutive csr entries of one
    // row (shell S, component x), summed in entry order into kseg; then each
    // row sums its jobs in job order.
    const int4* kjob;           // [nkjob] {S * HFK_MAXCOMP + x, first entry, end entry, -}
    int nkjob;
    double* kseg;               // [nkjob][ncols]
    const int4* krow;           // [nkrow] {S * HFK_MAXCOMP + x, first job, job count, -}
    int nkrow;
    const double* jab;          // [slot][NBRA] (bra side)
    const int* bcsr_off;        // [nbra + 1]: slots of each bra position
    const double* jpart;        // [nchunk][nket][NKET] ket-side chunk sums
    int nchunk;
    const int* bra_list;        // bra class list
    const int* ket_list;        // ket class list
    int nbra, nket, same;
    int na, nb, nc, nd;         // Cartesian components of the four shells
    double* Kpart;              // [n+1][n+1]
    double* Jpart;              // [n+1][n+1]
};
void hfk_launch_jk_part(const JKPart& P, int n, const int* sh_bf, const int* sh_bfp, const int* pair_A,
                        const int* pair_B, hipStream_t s);
// G = (J' + J'^T) - 1/2 (K' + K'^T) with J' = sum_e Jpart[e], K' = sum_e Kpart[e]
// (launches in plan order); G_flat = H_flat + G when H is given.
void hfk_launch_symmetrize_parts(int n, const double* Jpart, const double* Kpart, int nlaunch, const double* H,
                                 double* G_flat, hipStream_t s);
// flat -> padded (n+1)x(n+1) full, row/column n zero
void hfk_launch_flat_to_padded(int n, const double* flat, double* full, hipStream_t s);
// padded D ((n+1)^2, row/column n zero) from the flat density
void hfk_launch_fock_prepare(int n, const double* D_flat, double* D, hipStream_t s);
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
