/*
 * hfk_oracle.c -- CPU restatement of Haskell-abinitio's RHF integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * Every function follows a reference function; the citation is given on the
 * function (paths relative to the reference tree).  Arithmetic is written in
 * the same order as the reference where the order is observable at the
 * 1e-15 level (Boys table Taylor step, E and R recursions, prefactors).
 *
 * Parity status of each piece (what pins it) is listed in the header of
 * oracle/hfk_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define OB_BOYS_M 15      /* m = 0..14            TableBoys.hs:62 */
#define OB_BOYS_NX 1001   /* x = 0.1*i, i<=1000   TableBoys.hs:63 */

/* ------------------------------------------------------------------ */
/* Basis: one entry per reference CGF (one Cartesian component each).  */
/* ------------------------------------------------------------------ */
typedef struct {
    int nbf;
    const double *center;   /* [nbf][3]  bohr                           */
    const int *lxyz;        /* [nbf][3]  Cartesian exponents             */
    const int *prim_off;    /* [nbf+1]                                   */
    const double *coef;     /* [nprim]   as stored in CGF (normCoeff'd)  */
    const double *expo;     /* [nprim]                                   */
} ob_basis;

/* GHC's (^) for a non-negative Int exponent: repeated squaring in the
 * order of GHC.Real.(^) (f/g helper pair). */
static double hs_powi(double x, int n)
{
    if (n == 0) return 1.0;
    /* f x y */
    for (;;) {
        if ((n & 1) == 0) { x = x * x; n >>= 1; continue; }
        if (n == 1) return x;
        /* g (x*x) (y quot 2) x */
        double z = x;
        x = x * x; n >>= 1;
        for (;;) {
            if ((n & 1) == 0) { x = x * x; n >>= 1; continue; }
            if (n == 1) return x * z;
            z = x * z; x = x * x; n >>= 1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Gauss-Legendre 256 nodes: NumericalTools/PointsWeights.hs:25-31     */
/* (computed here by Newton iteration on P_256 in long double).        */
/* ------------------------------------------------------------------ */
static double gl_x[128], gl_w[128];
static int gl_ready = 0;

static void gl256_init(void)
{
    if (gl_ready) return;
    const int N = 256;
    for (int i = 0; i < N / 2; i++) {
        long double z = cosl((long double)M_PI * (i + 0.75L) / (N + 0.5L));
        long double pp = 0.0L;
        for (int it = 0; it < 100; it++) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= N; j++) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = N * (z * p1 - p2) / (z * z - 1.0L);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) < 1e-19L) break;
        }
        /* recompute derivative at the converged root */
        {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= N; j++) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = N * (z * p1 - p2) / (z * z - 1.0L);
        }
        /* i=0 is the largest root; store ascending */
        gl_x[N / 2 - 1 - i] = (double)z;
        gl_w[N / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
    }
    gl_ready = 1;
}

void ob_gl256(double *x, double *w)
{
    gl256_init();
    memcpy(x, gl_x, sizeof gl_x);
    memcpy(w, gl_w, sizeof gl_w);
}

/* fac with Int (64-bit, wrapping) arithmetic: Boys.hs:97-100 */
static int64_t hs_fac(int i)
{
    uint64_t acc = 1;
    for (int k = 1; k <= i; k++) acc *= (uint64_t)k;
    return (int64_t)acc;
}

/* facOdd: Boys.hs:89-94 */
static double hs_facodd(int i)
{
    if (i < 2) return 1.0;             /* compare i 2 == LT (odd i only) */
    int k = (1 + i) / 2;
    return (double)hs_fac(2 * k) / (hs_powi(2.0, k) * (double)hs_fac(k));
}

/* asymptBoysF: Boys.hs:85-87 */
double ob_boys_asympt(double m, double x)
{
    double fn = hs_facodd((int)floor(2.0 * m - 1.0));
    return (fn / pow(2.0, m + 1.0)) * sqrt(M_PI / pow(x, 2.0 * m + 1.0));
}

/* boysF: Boys.hs:54-66 with kummer Boys.hs:106-118 (err=0.1 selects
 * nIntegrate256, Boys.hs:155-157,163-165) */
double ob_boys_node(double m, double x)
{
    if (m == 0.0) {
        if (x > 50.0) return ob_boys_asympt(0.0, x);
        if (x > 0.0) return 0.5 * sqrt(M_PI / x) * erf(sqrt(x));
        return 1.0;
    }
    gl256_init();
    double a = m + 0.5, b = m + 1.5, z = -x;
    double gammaFactor = tgamma(b) / (tgamma(a) * tgamma(b - a));
    double e = exp(1.0);
    double s = 0.0;
    for (int i = 0; i < 128; i++) {
        double xp = 0.5 * ((1.0 * gl_x[i]) + 1.0 + 0.0);
        double xm = 0.5 * ((1.0 * -gl_x[i]) + 1.0 + 0.0);
        double fp = pow(e, z * xp) * pow(1.0 - xp, b - a - 1.0) * pow(xp, a - 1.0);
        double fm = pow(e, z * xm) * pow(1.0 - xm, b - a - 1.0) * pow(xm, a - 1.0);
        s += gl_w[i] * (fp + fm);
    }
    double integral = 0.5 * (1.0 - 0.0) * s;
    double k = gammaFactor * integral;
    return k / ((2.0 * m) + 1.0);
}

/* generateGridBoys 0.1: TableBoys.hs:58-63.  tab[m*1001+i] = boysF m (0.1*i) */
void ob_boys_table(double *tab)
{
    for (int m = 0; m < OB_BOYS_M; m++)
        for (int i = 0; i < OB_BOYS_NX; i++)
            tab[m * OB_BOYS_NX + i] = ob_boys_node((double)m, 0.1 * (double)i);
}

/* calcBoys: TableBoys.hs:65-69 */
static double boys_lookup(const double *tab, int m, double xk, long idx)
{
    if (xk > 100.0) return ob_boys_asympt((double)m, xk);
    if (m >= OB_BOYS_M || idx < 0 || idx >= OB_BOYS_NX) return NAN; /* reference: error */
    return tab[m * OB_BOYS_NX + idx];
}

/* boysTaylor: TableBoys.hs:45-51 */
double ob_boys_taylor(const double *tab, int m, double x)
{
    const double delta = 0.1;
    double fl = floor(x / delta);
    double xk = fl * delta;
    long idx = (long)fl;
    double dx = x - xk;
    double f0 = boys_lookup(tab, m, xk, idx);
    double f1 = boys_lookup(tab, m + 1, xk, idx);
    double f2 = boys_lookup(tab, m + 2, xk, idx);
    double f3 = boys_lookup(tab, m + 3, xk, idx);
    double f4 = boys_lookup(tab, m + 4, xk, idx);
    double f5 = boys_lookup(tab, m + 5, xk, idx);
    double r = f0 - f1 * dx;
    r = r + (0.5 * f2) * hs_powi(dx, 2);
    r = r - ((1.0 / 6.0) * f3) * hs_powi(dx, 3);
    r = r + ((1.0 / 24.0) * f4) * hs_powi(dx, 4);
    r = r - ((1.0 / 120.0) * f5) * hs_powi(dx, 5);
    return r;
}

/* Accurate Boys function (not the reference's): series for T < 50 + 2m,
 * closed-form F0 plus upward recursion above (stable there).  Used by the
 * ACCURATE mode. */
double ob_boys_exact(int m, double T)
{
    if (T < 50.0 + 2.0 * m) {
        long double term = 1.0L / (2.0L * m + 1.0L), s = term;
        for (int k = 1; k < 4000; k++) {
            term *= 2.0L * T / (2.0L * m + 2.0L * k + 1.0L);
            s += term;
            if (term < 1e-22L * s) break;
        }
        return (double)(expl(-(long double)T) * s);
    }
    long double F = 0.5L * sqrtl((long double)M_PI / T) * erfl(sqrtl((long double)T));
    long double et = expl(-(long double)T);
    for (int k = 0; k < m; k++) F = ((2.0L * k + 1.0L) * F - et) / (2.0L * T);
    return (double)F;
}

/* ------------------------------------------------------------------ */
/* Hermite (McMurchie-Davidson) machinery: IntegralsEvaluation.hs      */
/* ------------------------------------------------------------------ */

/* recursiveHermCoeff: IntegralsEvaluation.hs:457-477 */
static double herm_coeff(double gamma, double xpa, double xpb, double e0, int i, int j, int t)
{
    if (i == 0 && j == 0 && t == 0) return e0;
    if (i < 0 || j < 0 || t < 0) return 0.0;
    if (t == 0 && i >= 1) {
        double c0 = herm_coeff(gamma, xpa, xpb, e0, i - 1, j, 0);
        double c1 = herm_coeff(gamma, xpa, xpb, e0, i - 1, j, 1);
        return xpa * c0 + c1;
    }
    if (t == 0 && j >= 1) {
        double c0 = herm_coeff(gamma, xpa, xpb, e0, i, j - 1, 0);
        double c1 = herm_coeff(gamma, xpa, xpb, e0, i, j - 1, 1);
        return xpb * c0 + c1;
    }
    double a = herm_coeff(gamma, xpa, xpb, e0, i - 1, j, t - 1);
    double b = herm_coeff(gamma, xpa, xpb, e0, i, j - 1, t - 1);
    return (1.0 / (2.0 * gamma * (double)t)) * ((double)i * a + (double)j * b);
}

/* Hermite coefficient vector of a primitive pair in genCoeff_Hermite order
 * (t outer, v inner): calcHermCoeff + initilizedSeedCoeff,
 * IntegralsEvaluation.hs:452-455,489-502.  Returns the count. */
static int herm_vec(const double *ra, const double *rb, const int *la, const int *lb,
                    double ea, double eb, double mu, double gamma, const double *rp,
                    double *out, int *tuv)
{
    double e0[3], pa[3], pb[3];
    for (int x = 0; x < 3; x++) {
        double d = ra[x] - rb[x];
        e0[x] = exp(-(mu * (d * d)));
        pa[x] = rp[x] - ra[x];
        pb[x] = rp[x] - rb[x];
    }
    (void)ea; (void)eb;
    int n = 0;
    for (int t = 0; t <= la[0] + lb[0]; t++)
        for (int u = 0; u <= la[1] + lb[1]; u++)
            for (int v = 0; v <= la[2] + lb[2]; v++) {
                double cx = herm_coeff(gamma, pa[0], pb[0], e0[0], la[0], lb[0], t);
                double cy = herm_coeff(gamma, pa[1], pb[1], e0[1], la[1], lb[1], u);
                double cz = herm_coeff(gamma, pa[2], pb[2], e0[2], la[2], lb[2], v);
                out[n] = ((1.0 * cx) * cy) * cz;   /* product = foldl (*) 1 */
                if (tuv) { tuv[3 * n] = t; tuv[3 * n + 1] = u; tuv[3 * n + 2] = v; }
                n++;
            }
    return n;
}

/* meanp: IntegralsEvaluation.hs:549-550 */
static void meanp(double e1, double e2, const double *a, const double *b, double *p)
{
    for (int x = 0; x < 3; x++) p[x] = (e1 * a[x] + e2 * b[x]) / (e1 + e2);
}

/* Boys mode for the Hermite integrals: 0 = reference table, 1 = accurate */
typedef struct {
    const double *tab;
    int exact;
} ob_boys;

static double boys_eval(const ob_boys *bz, int m, double x)
{
    if (bz->exact) return ob_boys_exact(m, x);
    return ob_boys_taylor(bz->tab, m, x);
}

/* calcHermM: IntegralsEvaluation.hs:430-449, memoised over (n,t,u,v)
 * (the reference memoises with monad-memo; memoisation does not change
 * values).  d = L+1 bounds every index. */
#define RMAX 13
typedef struct {
    double x, y, z, alpha, arg;
    const ob_boys *bz;
    int d;
    double memo[RMAX * RMAX * RMAX * RMAX];
    unsigned char have[RMAX * RMAX * RMAX * RMAX];
} rstate;

static double herm_R(rstate *s, int n, int t, int u, int v)
{
    if (t < 0 || u < 0 || v < 0) return 0.0;
    int d = s->d;
    int key = ((n * d + t) * d + u) * d + v;
    if (s->have[key]) return s->memo[key];
    double r;
    if (t >= 1) {
        double b1 = herm_R(s, n + 1, t - 2, u, v);
        double b2 = herm_R(s, n + 1, t - 1, u, v);
        r = ((double)t - 1.0) * b1 + s->x * b2;
    } else if (u >= 1) {
        double b1 = herm_R(s, n + 1, t, u - 2, v);
        double b2 = herm_R(s, n + 1, t, u - 1, v);
        r = ((double)u - 1.0) * b1 + s->y * b2;
    } else if (v >= 1) {
        double b1 = herm_R(s, n + 1, t, u, v - 2);
        double b2 = herm_R(s, n + 1, t, u, v - 1);
        r = ((double)v - 1.0) * b1 + s->z * b2;
    } else {
        r = hs_powi(-2.0 * s->alpha, n) * boys_eval(s->bz, n, s->arg);
    }
    s->have[key] = 1;
    s->memo[key] = r;
    return r;
}

static void rstate_init(rstate *s, const double *rpq, double alpha, const ob_boys *bz, int L)
{
    s->x = rpq[0]; s->y = rpq[1]; s->z = rpq[2];
    s->alpha = alpha;
    s->arg = alpha * ((rpq[0] * rpq[0] + rpq[1] * rpq[1]) + rpq[2] * rpq[2]);
    s->bz = bz;
    s->d = L + 1;
    memset(s->have, 0, (size_t)(L + 1) * (L + 1) * (L + 1) * (L + 1));
}

static __thread rstate *tls_rs = NULL;
static rstate *get_rstate(void)
{
    if (!tls_rs) tls_rs = (rstate *)malloc(sizeof(rstate));
    return tls_rs;
}

/* twoTermsERI: IntegralsEvaluation.hs:319-337 (the reference screen value) */
double ob_two_terms(const double *ra, const int *la, double c1, double e1,
                    const double *rb, const int *lb, double c2, double e2)
{
    double coeff[64];
    int tuv[3 * 64];
    double p = e1 + e2;
    double rp[3];
    meanp(e1, e2, ra, rb, rp);
    double mu = e1 * e2 * (1.0 / (e1 + e2));
    int n = herm_vec(ra, rb, la, lb, e1, e2, mu, p, rp, coeff, tuv);
    double suma = 0.0;
    for (int k = 0; k < n; k++) {
        double sg = ((tuv[3 * k] + tuv[3 * k + 1] + tuv[3 * k + 2]) & 1) ? -1.0 : 1.0;
        suma += sg * coeff[k];
    }
    double acc = 0.0;
    for (int k = 0; k < n; k++) acc += coeff[k] * suma;
    double cte = 2.0 * ((c1 * c1) * (c2 * c2)) * sqrt(hs_powi(M_PI / (2.0 * p), 5));
    return cte * acc;
}

/* twoElectronHermite + mcMurchie2: IntegralsEvaluation.hs:340-376 */
static double prim_eri(const double *ra, const int *la, double c1, double e1,
                       const double *rb, const int *lb, double c2, double e2,
                       const double *rc, const int *lc, double c3, double e3,
                       const double *rd, const int *ld, double c4, double e4,
                       const ob_boys *bz, rstate *rs)
{
    double coeff1[128], coeff2[128];
    int tuv1[3 * 128], tuv2[3 * 128];
    double rp[3], rq[3], rpq[3];
    meanp(e1, e2, ra, rb, rp);
    meanp(e3, e4, rc, rd, rq);
    double mu = e1 * e2 * (1.0 / (e1 + e2));
    double nu = e3 * e4 * (1.0 / (e3 + e4));
    double p = e1 + e2, q = e3 + e4;
    double alpha = p * q / (p + q);
    for (int x = 0; x < 3; x++) rpq[x] = rp[x] - rq[x];
    double cte = (((c1 * c2) * c3) * c4) * ((1.0 / ((p * q) * sqrt(p + q))) * (2.0 * pow(M_PI, 2.5)));
    int n1 = herm_vec(ra, rb, la, lb, e1, e2, mu, p, rp, coeff1, tuv1);
    int n2 = herm_vec(rc, rd, lc, ld, e3, e4, nu, q, rq, coeff2, tuv2);
    rstate_init(rs, rpq, alpha, bz,
                la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2] + lc[0] + lc[1] + lc[2] + ld[0] + ld[1] + ld[2]);
    double s = 0.0;
    for (int a = 0; a < n1; a++) {
        double m = 0.0;
        for (int b = 0; b < n2; b++) {
            int tt = tuv2[3 * b], uu = tuv2[3 * b + 1], vv = tuv2[3 * b + 2];
            double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
            double r = herm_R(rs, 0, tuv1[3 * a] + tt, tuv1[3 * a + 1] + uu, tuv1[3 * a + 2] + vv);
            m += (sg * coeff2[b]) * r;
        }
        s += coeff1[a] * m;
    }
    return cte * s;
}

/* contracted4Centers + schwarz: IntegralsEvaluation.hs:294-315.
 * mode_screen: 1 = reference per-primitive screen (1e-8), 0 = none.
 * *nprim_kept counts primitive quartets that pass. */
double ob_eri(const ob_basis *bs, const double *tab, int exact, int screen,
              int i, int j, int k, int l, long *nprim_kept)
{
    ob_boys bz = {tab, exact};
    rstate *rs = get_rstate();
    const int idx[4] = {i, j, k, l};
    const double *r[4];
    const int *lx[4];
    for (int a = 0; a < 4; a++) { r[a] = bs->center + 3 * idx[a]; lx[a] = bs->lxyz + 3 * idx[a]; }
    double sum = 0.0;
    long kept = 0;
    for (int g1 = bs->prim_off[i]; g1 < bs->prim_off[i + 1]; g1++)
        for (int g2 = bs->prim_off[j]; g2 < bs->prim_off[j + 1]; g2++) {
            double qab = sqrt(ob_two_terms(r[0], lx[0], bs->coef[g1], bs->expo[g1],
                                           r[1], lx[1], bs->coef[g2], bs->expo[g2]));
            for (int g3 = bs->prim_off[k]; g3 < bs->prim_off[k + 1]; g3++)
                for (int g4 = bs->prim_off[l]; g4 < bs->prim_off[l + 1]; g4++) {
                    if (screen) {
                        double qcd = sqrt(ob_two_terms(r[2], lx[2], bs->coef[g3], bs->expo[g3],
                                                       r[3], lx[3], bs->coef[g4], bs->expo[g4]));
                        double val = qab * qcd;
                        if (val < 1e-8) continue;   /* guard (not (val < 1e-8)) */
                    }
                    kept++;
                    sum += prim_eri(r[0], lx[0], bs->coef[g1], bs->expo[g1],
                                    r[1], lx[1], bs->coef[g2], bs->expo[g2],
                                    r[2], lx[2], bs->coef[g3], bs->expo[g3],
                                    r[3], lx[3], bs->coef[g4], bs->expo[g4], &bz, rs);
                }
        }
    if (nprim_kept) *nprim_kept = kept;
    return sum;
}

/* ------------------------------------------------------------------ */
/* Layout maps: NumericalTools/LinearAlgebra.hs:240-284               */
/* ------------------------------------------------------------------ */
static long calc_elem(long dim, long w)
{
    /* sum . takeZero w $ iterate pred dim  =  sum_{t<w} (dim - t) */
    if (w <= 0) return 0;
    return w * dim - w * (w - 1) / 2;
}

/* indexDIM2toFlat: LinearAlgebra.hs:250-253 */
long ob_index_dim2_to_flat(long dim, long x, long y)
{
    if (x <= y) return calc_elem(dim, x) + (y - x);
    return calc_elem(dim, y) + (x - y);
}

static long sum_index_triang(long n) { return (n * n + n) / 2; }

static long sum_first_index(long n)
{
    long s = n + sum_index_triang(n - 1);
    for (long k = 1; k <= n - 1; k++) s += sum_index_triang(n) - k;
    return s;
}

/* fourIndex2Flat: LinearAlgebra.hs:266-283, restated literally
 * (including its behaviour for out-of-range indices). */
long ob_four_index_ref(long n, long i, long j, long k, long l)
{
    long a0 = 0;
    for (long m = 0; m <= i - 1; m++) a0 += sum_first_index(n - m);
    long pos;
    if (i == j && i == k) {
        pos = l - i;
    } else if (i == j) {
        long ip = i + 1;
        pos = (n - i) + ob_index_dim2_to_flat(n - ip, k - ip, l - ip);
    } else {
        long s = 0;
        for (long m = 1; m <= (j - 1) - i; m++) s += sum_index_triang(n - i) - m;
        pos = (n - i) + sum_index_triang((n - 1) - i) + s
              + (ob_index_dim2_to_flat(n, k, l) - ob_index_dim2_to_flat(n, i, j));
    }
    return a0 + pos;
}

/* sortKeys: HartreeFock.hs:153-156 */
void ob_sort_keys(const long *in, long *out)
{
    long i = in[0], j = in[1], k = in[2], l = in[3];
    long a0 = i < j ? i : j, a1 = i < j ? j : i;
    long b0 = k < l ? k : l, b1 = k < l ? l : k;
    int le = (a0 < b0) || (a0 == b0 && a1 <= b1);
    if (le) { out[0] = a0; out[1] = a1; out[2] = b0; out[3] = b1; }
    else { out[0] = b0; out[1] = b1; out[2] = a0; out[3] = a1; }
}

/* Closed-form rank of a canonical quartet in calcIntegrals' enumeration
 * order (HartreeFock.hs:166-181): i-major, then j, k, l. */
long ob_flat_rank(long n, long i, long j, long k, long l)
{
    /* count(i',j') = T(n-1-i') + (n-j') tuples share (i',j') */
    long r = 0;
    for (long ii = 0; ii < i; ii++)
        for (long jj = ii; jj < n; jj++) r += (n - 1 - ii) * (n - ii) / 2 + (n - jj);
    for (long jj = i; jj < j; jj++) r += (n - 1 - i) * (n - i) / 2 + (n - jj);
    if (k == i) return r + (l - j);
    r += n - j;
    for (long kk = i + 1; kk < k; kk++) r += n - kk;
    return r + (l - k);
}

long ob_flat_len(long n)
{
    long m = n * (n + 1) / 2;
    return m * (m + 1) / 2;
}

/* calcIntegrals: HartreeFock.hs:166-181.  out has ob_flat_len(nbf) entries,
 * written in enumeration order.  Parallelised over i with OpenMP. */
void ob_eri_flat(const ob_basis *bs, const double *tab, int exact, int screen, double *out,
                 long *nprim_kept)
{
    long n = bs->nbf;
    long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (long i = 0; i < n; i++) {
        long pos = ob_flat_rank(n, i, i, i, i);
        for (long j = i; j < n; j++)
            for (long k = i; k < n; k++)
                for (long l = k; l < n; l++) {
                    if (!(k > i || (k == i && l >= j))) continue;
                    long kept = 0;
                    out[pos++] = ob_eri(bs, tab, exact, screen, (int)i, (int)j, (int)k, (int)l, &kept);
                    total += kept;
                }
    }
    if (nprim_kept) *nprim_kept = total;
}

/* base[i*n+j] = rank of (i,j,i,j); with it a canonical rank is O(1). */
static long *rank_base(long n)
{
    long *base = (long *)malloc(sizeof(long) * (size_t)n * (size_t)n);
    long r = 0;
    for (long i = 0; i < n; i++)
        for (long j = i; j < n; j++) {
            base[i * n + j] = r;
            r += (n - 1 - i) * (n - i) / 2 + (n - j);
        }
    return base;
}

static long rank_fast(const long *base, long n, long i, long j, long k, long l)
{
    long r = base[i * n + j];
    if (k == i) return r + (l - j);
    r += n - j;
    long cnt = k - 1 - i;
    r += cnt * n - ((k - 1) * k / 2 - i * (i + 1) / 2);
    return r + (l - k);
}

/* Index of the reference-sorted quartet in the flat array. */
static long flat_of(long n, long i, long j, long k, long l)
{
    long in[4] = {i, j, k, l}, s[4];
    ob_sort_keys(in, s);
    return ob_flat_rank(n, s[0], s[1], s[2], s[3]);
}

static long flat_of_fast(const long *base, long n, long i, long j, long k, long l)
{
    long in[4] = {i, j, k, l}, s[4];
    ob_sort_keys(in, s);
    return rank_fast(base, n, s[0], s[1], s[2], s[3]);
}

/* calcGmatrix + sliceIntegrals: HartreeFock.hs:225-260.
 * D_flat, G_flat: upper-triangular row-major (LinearAlgebra.hs:240-253). */
void ob_gmatrix(long n, const double *eri, const double *D_flat, double *G_flat)
{
    long flat = n * (n + 1) / 2;
    long *base = rank_base(n);
#pragma omp parallel for schedule(dynamic, 16)
    for (long idx = 0; idx < flat; idx++) {
        /* indexFlat2DIM2 */
        long a = 0, rem = idx;
        while (rem >= n - a) { rem -= n - a; a++; }
        long b = a + rem;
        double outer = 0.0;
        for (long l = 0; l < n; l++) {
            double inner = 0.0;
            for (long k = 0; k < n; k++) {
                double dlk = D_flat[ob_index_dim2_to_flat(n, l, k)];
                double coul = eri[flat_of_fast(base, n, a, b, k, l)];
                double exch = eri[flat_of_fast(base, n, a, l, k, b)];
                inner += dlk * (coul - 0.5 * exch);
            }
            outer += inner;
        }
        G_flat[idx] = outer;
    }
    free(base);
}

/* Reference-sorted lookup of one ERI from a flat array (tests). */
double ob_flat_get(long n, const double *eri, long i, long j, long k, long l)
{
    return eri[flat_of(n, i, j, k, l)];
}

/* ------------------------------------------------------------------ */
/* One-electron integrals: IntegralsEvaluation.hs:169-280,385-414      */
/* ------------------------------------------------------------------ */

/* obaraSaika: IntegralsEvaluation.hs:209-221 */
static double os1d(double gamma, double s00, double pax, double pbx, int m, int n)
{
    double c = 1.0 / (2.0 * gamma);
    if (m < 0 || n < 0) return 0.0;
    if (m == 0 && n == 0) return s00;
    double m1 = (double)(m - 1), n1 = (double)(n - 1);
    if (m == 0) return pbx * os1d(gamma, s00, pax, pbx, 0, n - 1) + c * n1 * os1d(gamma, s00, pax, pbx, 0, n - 2);
    if (n == 0) return pax * os1d(gamma, s00, pax, pbx, m - 1, 0) + c * m1 * os1d(gamma, s00, pax, pbx, m - 2, 0);
    return pax * os1d(gamma, s00, pax, pbx, m - 1, n)
           + c * (m1 * os1d(gamma, s00, pax, pbx, m - 2, n) + (double)n * os1d(gamma, s00, pax, pbx, m - 1, n - 1));
}

typedef struct { double gamma, s00[3], pa[3], pb[3]; } os_pair;

static void os_pair_init(os_pair *o, const double *r1, double e1, const double *r2, double e2)
{
    double p[3];
    meanp(e1, e2, r1, r2, p);
    double gamma = e1 + e2, mu = e1 * e2 / gamma, cte = sqrt(M_PI / gamma);
    o->gamma = gamma;
    for (int x = 0; x < 3; x++) {
        double d = r1[x] - r2[x];
        o->s00[x] = cte * exp(-(mu * (d * d)));
        o->pa[x] = p[x] - r1[x];
        o->pb[x] = p[x] - r2[x];
    }
}

/* sab: IntegralsEvaluation.hs:195-206 */
static double prim_overlap(const double *r1, const int *l1, double c1, double e1,
                           const double *r2, const int *l2, double c2, double e2)
{
    os_pair o;
    os_pair_init(&o, r1, e1, r2, e2);
    double pr = 1.0;
    for (int x = 0; x < 3; x++) pr *= os1d(o.gamma, o.s00[x], o.pa[x], o.pb[x], l1[x], l2[x]);
    return (c1 * c2) * pr;
}

/* tab: IntegralsEvaluation.hs:263-280 */
static double prim_kinetic(const double *r1, const int *l1, double c1, double e1,
                           const double *r2, const int *l2, double c2, double e2)
{
    os_pair o;
    os_pair_init(&o, r1, e1, r2, e2);
    double sx[3], tx[3];
    for (int x = 0; x < 3; x++) {
        double L2 = (double)l2[x];
        sx[x] = os1d(o.gamma, o.s00[x], o.pa[x], o.pb[x], l1[x], l2[x]);
        tx[x] = -2.0 * hs_powi(e2, 2) * os1d(o.gamma, o.s00[x], o.pa[x], o.pb[x], l1[x], l2[x] + 2)
                + e2 * (2.0 * L2 + 1.0) * sx[x]
                - 0.5 * L2 * (L2 - 1.0) * os1d(o.gamma, o.s00[x], o.pa[x], o.pb[x], l1[x], l2[x] - 2);
    }
    double t1 = ((1.0 * tx[0]) * sx[1]) * sx[2];
    double t2 = ((1.0 * sx[0]) * tx[1]) * sx[2];
    double t3 = ((1.0 * sx[0]) * sx[1]) * tx[2];
    return c1 * c2 * ((t1 + t2) + t3);
}

/* vijHermite + mcMurchie: IntegralsEvaluation.hs:395-414 (no derivative) */
static double prim_nuclear(const double *ra, const int *la, double c1, double e1,
                           const double *rb, const int *lb, double c2, double e2,
                           const double *rc, const ob_boys *bz, rstate *rs)
{
    double coeff[128];
    int tuv[3 * 128];
    double gamma = e1 + e2, nu = e1 * e2 / gamma;
    double rp[3], rpc[3];
    meanp(e1, e2, ra, rb, rp);
    for (int x = 0; x < 3; x++) rpc[x] = rp[x] - rc[x];
    int n = herm_vec(ra, rb, la, lb, e1, e2, nu, gamma, rp, coeff, tuv);
    rstate_init(rs, rpc, gamma, bz, la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2]);
    double s = 0.0;
    for (int k = 0; k < n; k++) s += coeff[k] * herm_R(rs, 0, tuv[3 * k], tuv[3 * k + 1], tuv[3 * k + 2]);
    double cte = c1 * c2 * 2.0 * (M_PI / gamma);
    return 1.0 * cte * s;
}

/* mtxOverlap + sijContracted (S_ii := 1): IntegralsEvaluation.hs:169-191.
 * natom/atom_xyz are not needed; S_flat has nbf(nbf+1)/2 entries. */
void ob_overlap(const ob_basis *bs, double *S_flat)
{
    long n = bs->nbf, idx = 0;
    for (long i = 0; i < n; i++)
        for (long j = i; j < n; j++, idx++) {
            if (i == j) { S_flat[idx] = 1.0; continue; }
            double s = 0.0;
            for (int g1 = bs->prim_off[i]; g1 < bs->prim_off[i + 1]; g1++)
                for (int g2 = bs->prim_off[j]; g2 < bs->prim_off[j + 1]; g2++)
                    s += prim_overlap(bs->center + 3 * i, bs->lxyz + 3 * i, bs->coef[g1], bs->expo[g1],
                                      bs->center + 3 * j, bs->lxyz + 3 * j, bs->coef[g2], bs->expo[g2]);
            S_flat[idx] = s;
        }
}

/* hcore: IntegralsEvaluation.hs:227-240 (T + sum_C (-Z_C) V_C) */
void ob_hcore(const ob_basis *bs, const double *tab, int exact, int natom, const double *atom_xyz,
              const double *Z, double *H_flat)
{
    ob_boys bz = {tab, exact};
    long n = bs->nbf;
    long flat = n * (n + 1) / 2;
#pragma omp parallel
    {
        rstate *rs = get_rstate();
#pragma omp for schedule(dynamic, 8)
        for (long idx = 0; idx < flat; idx++) {
            long i = 0, rem = idx;
            while (rem >= n - i) { rem -= n - i; i++; }
            long j = i + rem;
            const double *ri = bs->center + 3 * i, *rj = bs->center + 3 * j;
            const int *li = bs->lxyz + 3 * i, *lj = bs->lxyz + 3 * j;
            double t = 0.0;
            for (int g1 = bs->prim_off[i]; g1 < bs->prim_off[i + 1]; g1++)
                for (int g2 = bs->prim_off[j]; g2 < bs->prim_off[j + 1]; g2++)
                    t += prim_kinetic(ri, li, bs->coef[g1], bs->expo[g1], rj, lj, bs->coef[g2], bs->expo[g2]);
            double sv = 0.0;
            for (int c = 0; c < natom; c++) {
                double v = 0.0;
                for (int g1 = bs->prim_off[i]; g1 < bs->prim_off[i + 1]; g1++)
                    for (int g2 = bs->prim_off[j]; g2 < bs->prim_off[j + 1]; g2++)
                        v += prim_nuclear(ri, li, bs->coef[g1], bs->expo[g1], rj, lj, bs->coef[g2],
                                          bs->expo[g2], atom_xyz + 3 * c, &bz, rs);
                sv += (-Z[c]) * v;
            }
            H_flat[idx] = t + sv;
        }
    }
}

/* A list of function quartets (4 indices each), evaluated with the
 * reference algorithm (contracted4Centers) in parallel: the CPU baseline. */
void ob_eri_list(const ob_basis *bs, const double *tab, int exact, int screen, long nq, const int *q, double *out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long t = 0; t < nq; t++)
        out[t] = ob_eri(bs, tab, exact, screen, q[4 * t], q[4 * t + 1], q[4 * t + 2], q[4 * t + 3], NULL);
}
