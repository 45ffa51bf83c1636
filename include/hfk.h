/*
 * hfk.h -- C ABI of the MI355X restricted Hartree-Fock integral engine.
 *
 * Drop-in replacement for the Repa-facing functions of Haskell-abinitio's
 * RHF hot path.  Each entry point names the reference function it replaces
 * (paths relative to the reference tree).  Plain pointers and sizes only.
 *
 * Layouts (identical to the reference):
 *  - "flat" symmetric matrices: upper triangle, row-major, diagonal included,
 *    n(n+1)/2 entries; element (i,j), i<=j, at sum_{t<i}(n-t) + (j-i)
 *    (NumericalTools/LinearAlgebra.hs:240-253);
 *  - square matrices: n*n row-major (Repa DIM2 U arrays);
 *  - the flat ERI array: calcIntegrals' enumeration order, i<=j, k<=l,
 *    (i,j) <= (k,l) lexicographically, i-major then j, k, l
 *    (HartreeFock/HartreeFock.hs:166-181), i.e. fourIndex2Flat order
 *    (LinearAlgebra.hs:266-283).
 *
 * Ownership: the caller allocates every output buffer; the library never
 * frees caller memory.  Host-pointer entry points copy in and out; the *_dev
 * entry points take device pointers (memory from hipMalloc on the context's
 * device) and leave results in HBM.
 *
 * Errors: every function returns 0 on success and a negative HFK_E* code
 * otherwise; hfk_last_error() gives the message of the last failure on the
 * calling thread.  Nothing throws or longjmps across this boundary.
 *
 * Threading: a context serialises its own calls (one mutex, one HIP
 * stream); distinct contexts may be used from distinct threads.
 */
#ifndef HFK_H
#define HFK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFK_OK 0
#define HFK_EINVAL -1      /* bad argument                        */
#define HFK_EHIP -2        /* HIP / rocBLAS / rocSOLVER failure    */
#define HFK_ENOMEM -3      /* device allocation failed             */
#define HFK_ESCF -4        /* SCF did not converge (maxStep)       */
#define HFK_EUNSUP -5      /* unsupported input (e.g. l > 2)       */
#define HFK_ENUMERIC -6    /* numerical failure (NaN, singular S)  */

/* One entry per reference CGF (one Cartesian component), in [AtomData]
 * order and getBasis order (GlobalTypes.hs:92-105). */
typedef struct hfk_basis {
    int32_t nbf;
    int32_t nprim_total;
    int32_t natom;
    const double *center;      /* [nbf][3] bohr (AtomData.getCoord)                  */
    const int32_t *lxyz;       /* [nbf][3] Cartesian exponents (mapAxesMomentum)     */
    const int32_t *prim_off;   /* [nbf+1]                                            */
    const double *coef;        /* [nprim] as stored in the CGF (normCoeff-scaled)    */
    const double *expo;        /* [nprim]                                            */
    const int32_t *atom_of_bf; /* [nbf]                                              */
    const double *Z;           /* [natom] nuclear charges (getZnumber)               */
    const double *R;           /* [natom][3] nuclear positions, bohr                 */
} hfk_basis;

#define HFK_SCREEN_REFERENCE 0 /* keep a primitive quartet iff !(q_ab q_cd < 1e-8),
                                  per Cartesian component (IntegralsEvaluation.hs:305-315) */
#define HFK_SCREEN_NONE 1

#define HFK_BOYS_REFERENCE 0   /* 0.1-grid table, 6-term Taylor (TableBoys.hs:45-69) */
#define HFK_BOYS_ACCURATE 1    /* accurate table + recursion (not the reference's)   */

typedef struct hfk_options {
    int32_t screen_mode;       /* HFK_SCREEN_*                                       */
    int32_t boys_mode;         /* HFK_BOYS_*                                         */
    double screen_tau;         /* primitive screen threshold (reference: 1e-8)      */
    double schwarz_cut;        /* skip shell quartets with Schwarz bound below this
                                  (0 = off, keeps exact reference semantics)         */
    double fp32_threshold;     /* shell quartets with Schwarz bound < threshold are
                                  evaluated and contracted in FP32 (0 = all FP64)    */
} hfk_options;

/* Fill *opt with the reference-compatible defaults: reference screen at
 * 1e-8, reference Boys table, no Schwarz cut, all FP64. */
void hfk_default_options(hfk_options *opt);

typedef struct hfk_ctx hfk_ctx;

/* Create a context on the current HIP device: uploads the basis, groups
 * CGFs into shells, builds primitive-pair data, the Boys tables and the
 * Schwarz bounds.  New (the reference has no context). */
int hfk_create(const hfk_basis *basis, const hfk_options *opt, hfk_ctx **out);
int hfk_destroy(hfk_ctx *ctx);
int hfk_set_options(hfk_ctx *ctx, const hfk_options *opt);

/* calcIntegrals :: Map Boys Double -> [AtomData] -> Array U DIM1 Double
 * (HartreeFock/HartreeFock.hs:166-181).  out: nbf(nbf+1)/2*(nbf(nbf+1)/2+1)/2
 * doubles in calcIntegrals order.  Meant for drop-in parity at small nbf. */
int hfk_eri_flat(hfk_ctx *ctx, double *out, size_t out_len);

/* No counterpart in the reference: hfk_eri_flat for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same calcIntegrals array,
 * written to out_dev (device memory of the context's device, out_len =
 * M(M+1)/2 doubles, M = nbf(nbf+1)/2).  The call allocates nothing: its one
 * scratch table lives in a fixed device array of the library, so nbf is at
 * most HFK_NOALLOC_MAX_NBF (HFK_EINVAL above).  All calls on one GPU share
 * that array: calls that overlap on one GPU race.  Runs on the context's
 * stream (work already queued for out_dev elsewhere must have completed) and
 * returns when the array is written. */
#define HFK_NOALLOC_MAX_NBF 256
int hfk_eri_flat_noalloc(hfk_ctx *ctx, double *out_dev, size_t out_len);

/* calcGmatrix (HartreeFock.hs:225-241): G_ab = sum_kl D_lk [(ab|kl) - 1/2 (al|kb)],
 * integral-direct.  D_flat, G_flat: flat, nbf(nbf+1)/2. */
int hfk_fock_G(hfk_ctx *ctx, const double *D_flat, double *G_flat);

/* fock (HartreeFock.hs:213-222): F = H + G. */
int hfk_fock(hfk_ctx *ctx, const double *H_flat, const double *D_flat, double *F_flat);

/* Device-resident Fock build: D_dev, G_dev are flat (nbf(nbf+1)/2) device
 * arrays; H_dev may be NULL (then G only).  Synchronous. */
int hfk_fock_dev(hfk_ctx *ctx, const double *H_dev, const double *D_dev, double *F_dev);

/* calcDensity (HartreeFock.hs:195-208): D_kl = 2 sum_{i<occ} C_ki C_li.
 * C: nbf x nbf row-major, MO coefficients in columns. */
int hfk_density(hfk_ctx *ctx, const double *C, int32_t occ, double *D_flat);

/* symmOrtho (HartreeFock/BasisOrthogonalization.hs:26-35): X = U s^-1/2 U^T.
 * S, X: nbf x nbf row-major. */
int hfk_symm_ortho(hfk_ctx *ctx, const double *S, double *X);

/* unitaryTransf (LinearAlgebra.hs:155-159): Fp = X^T F X.  X, Fp: nbf x nbf
 * row-major; F_flat flat. */
int hfk_unitary(hfk_ctx *ctx, const double *X, const double *F_flat, double *Fp);

/* mtxOverlap (IntegralsEvaluation.hs:169-177), S_ii := 1.  S_flat flat. */
int hfk_overlap(hfk_ctx *ctx, double *S_flat);

/* hcore (IntegralsEvaluation.hs:227-240): T + sum_C (-Z_C) V_C.  H_flat flat. */
int hfk_hcore(hfk_ctx *ctx, double *H_flat);

#define HFK_GUESS_CORE 0       /* D from diagonalising H                               */
#define HFK_GUESS_DENSITY 2    /* new: start from hfk_scf_options.guess_density        */
#define HFK_GUESS_INTENDED 1   /* Harris step with D = 2 on the first occ diagonal AOs,
                                  correctly indexed (HartreeFock.hs:128-149 intent)     */
#define HFK_GUESS_REFERENCE 3  /* harrisFunctional as written (HartreeFock.hs:127-149):
                                  flat length as the matrix dimension, occupied as the
                                  basis size; HFK_EINVAL where the reference's bounds-
                                  checked indexing aborts (from (H2O)_8 up)             */

typedef struct hfk_scf_options {
    int32_t guess;             /* HFK_GUESS_*                                         */
    int32_t max_steps;         /* reference: 30 (HartreeFock.hs:87)                    */
    int32_t diis_size;         /* reference: 5 (HartreeFock.hs:78)                     */
    double diis_start;         /* reference: 0.1 (DIIS.hs:58)                          */
    double converge;           /* reference: max|X^T(FDS-SDF)X| < 1e-7 (HartreeFock.hs:113) */
    double charge;
    int32_t verbose;           /* new: 1 = one progress line per step on stderr        */
    int32_t pad;
    const double *guess_density; /* new: flat starting density for HFK_GUESS_DENSITY     */
    double incremental_tau;    /* new: 0 = full Fock build every step (the reference);
                                  > 0: F_k = F(D_k-1) + G(D_k - D_k-1), skipping shell
                                  quartets with Q_ab Q_cd max|dD| < tau (Q: Schwarz)   */
    int32_t rebuild_every;     /* new: with incremental_tau, a full build every this
                                  many steps (<= 1: every step)                        */
    int32_t pad2;
    double energy_converge;    /* new: 0 = the reference's rule alone; > 0: stop only
                                  when also |E_k - E_{k-1}| < energy_converge          */
} hfk_scf_options;

typedef struct hfk_scf_result {
    double energy;             /* total energy incl. nuclear repulsion (variationalE)  */
    double e_nuc;
    double max_error;          /* final max |X^T(FDS-SDF)X|                            */
    int32_t steps;
    int32_t occupied;
    double seconds;            /* wall time of the SCF call                             */
    double fock_seconds;       /* time in Fock builds                                   */
    double delta_energy;       /* new: E_k - E_{k-1} of the last step k (+inf after a
                                  single step); for an energy-change criterion beside
                                  the reference's max|e| rule                          */
} hfk_scf_result;

void hfk_default_scf_options(hfk_scf_options *o);

/* scfHF (HartreeFock.hs:70-123): guess, DIIS loop, FP64 eigen-solve on the
 * GPU.  Any of C (nbf x nbf), eps (nbf), D_flat, F_flat may be NULL. */
int hfk_scf(hfk_ctx *ctx, const hfk_scf_options *o, hfk_scf_result *res, double *C, double *eps,
            double *D_flat, double *F_flat);

/* Nuclear gradient dE/dR for a converged density (new: the reference's
 * Derivatives.hs sketch does not compile).  D_flat: density; W_flat:
 * energy-weighted density 2 sum_{i<occ} eps_i C_ui C_vi.  g: [natom][3]. */
int hfk_gradient(hfk_ctx *ctx, const double *D_flat, const double *W_flat, double *g);

typedef struct hfk_stats {
    int64_t nshell;
    int64_t nshell_pair;          /* significant shell pairs kept                        */
    int64_t nprim_pair;
    int64_t quartets_launched;    /* shell quartets of the last Fock build (after pair screen) */
    int64_t func_quartets;        /* canonical function quartets they hold               */
    double last_fock_ms;          /* wall time of the last Fock build (ms, HIP events)   */
    double last_eri_ms;           /* ERI+digestion kernel time of the last Fock build    */
    double last_gemm_ms;          /* MFMA GEMM time of the last hfk_unitary / hfk_density call (HIP events) */
    double last_gemm_flops;       /* its floating-point operations (2mnk per GEMM)       */
} hfk_stats;

int hfk_get_stats(hfk_ctx *ctx, hfk_stats *st);

/* One record per kernel launch of a Fock build: angular-momentum class,
 * precision, counted survivors and the duration of its last run (HIP events
 * on the context's stream).  New (measurement; no reference counterpart). */
typedef struct hfk_class_stat {
    int32_t la, lb, lc, ld;
    int32_t prec;                 /* 0 FP64, 1 FP32                                     */
    int32_t solo;                 /* 1: its last run had the GPU to itself (ms is its own) */
    int64_t shell_quartets;       /* shell quartets passing the pair screens            */
    int64_t prim_quartets;        /* primitive quartets not screened as a whole         */
    int64_t func_quartets;        /* canonical function quartets they hold              */
    double ms;                    /* kernel time of the last Fock build / export        */
    int64_t tasks;                /* workgroups of its J/K launch (bra pair x ket run)   */
    double grad_ms;               /* its derivative-integral launch in the last hfk_gradient
                                     (FP64 records; 0 for the FP32 part)                  */
} hfk_class_stat;

int hfk_get_class_stats(hfk_ctx *ctx, hfk_class_stat *out, int32_t cap, int32_t *n);

int hfk_last_error(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* HFK_H */
