/*
 * sanitize_check.c -- runs every C entry point of the CPU oracle once under
 * AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
 * `sanitize`, run by tests/test_oracle.py).
 *
 * TEST INFRASTRUCTURE ONLY, like the oracle it exercises.  The file includes
 * hfk_oracle.c so that its static helpers are instrumented in the same unit.
 * It is built without OpenMP (the pragmas are ignored), so the thread-local
 * recursion state and every loop run on one thread.
 *
 * Besides running clean under the sanitizers, the run checks properties that
 * the mathematics fixes (exit status 1 on the first failure):
 *   - F_m(0) = 1/(2m+1) at the table's first node;
 *   - the table + Taylor step agrees with the accurate Boys function to the
 *     reference table's documented node error, and stays finite past x = 100;
 *   - the 8-fold permutation symmetry of contracted ERIs over s, p, d;
 *   - the flat layout: the literal fourIndex2Flat equals the enumeration rank
 *     of every canonical quartet, and ob_flat_get returns the canonical entry
 *     for every index permutation;
 *   - G from the flat ERIs equals the dense contraction of the same ERIs;
 *   - S and H are finite, with S_ii = 1.
 */
#include <stdio.h>

#include "hfk_oracle.c"

static int fails = 0;
#define CHECK(c, ...)                                 \
    do {                                              \
        if (!(c)) {                                   \
            fprintf(stderr, "FAIL: " __VA_ARGS__);    \
            fprintf(stderr, "\n");                    \
            fails++;                                  \
        }                                             \
    } while (0)

/* three centres: an s (3 primitives), p (2 primitives) and d (1 primitive)
 * shell on the first, an s (2 primitives) shell on each of the others */
enum { NBF = 12, NPRIM = 3 + 3 * 2 + 6 * 1 + 2 + 2 };

static double center[NBF * 3];
static int lxyz[NBF * 3];
static int prim_off[NBF + 1];
static double coef[NPRIM], expo[NPRIM];

static void make_basis(ob_basis *bs)
{
    static const double cen[3][3] = {{0.0, 0.0, 0.1}, {1.43, 0.0, -0.9}, {-1.43, 0.2, -0.9}};
    static const int pcomp[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static const int dcomp[6][3] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
    int f = 0, p = 0;
    /* O-like s */
    memcpy(center + 3 * f, cen[0], sizeof cen[0]);
    lxyz[3 * f] = lxyz[3 * f + 1] = lxyz[3 * f + 2] = 0;
    prim_off[f] = p;
    expo[p] = 130.7; coef[p++] = 0.154;
    expo[p] = 23.8; coef[p++] = 0.535;
    expo[p] = 6.44; coef[p++] = 0.445;
    f++;
    for (int c = 0; c < 3; c++, f++) {
        memcpy(center + 3 * f, cen[0], sizeof cen[0]);
        memcpy(lxyz + 3 * f, pcomp[c], sizeof pcomp[c]);
        prim_off[f] = p;
        expo[p] = 5.03; coef[p++] = 0.156;
        expo[p] = 1.17; coef[p++] = 0.607;
    }
    for (int c = 0; c < 6; c++, f++) {
        memcpy(center + 3 * f, cen[0], sizeof cen[0]);
        memcpy(lxyz + 3 * f, dcomp[c], sizeof dcomp[c]);
        prim_off[f] = p;
        expo[p] = 0.8; coef[p++] = 1.0;
    }
    for (int h = 1; h <= 2; h++, f++) {
        memcpy(center + 3 * f, cen[h], sizeof cen[h]);
        lxyz[3 * f] = lxyz[3 * f + 1] = lxyz[3 * f + 2] = 0;
        prim_off[f] = p;
        expo[p] = 3.43; coef[p++] = 0.154;
        expo[p] = 0.624; coef[p++] = 0.535;
    }
    prim_off[f] = p;
    bs->nbf = NBF;
    bs->center = center;
    bs->lxyz = lxyz;
    bs->prim_off = prim_off;
    bs->coef = coef;
    bs->expo = expo;
}

static int close_rel(double a, double b, double rel, double abs_)
{
    return fabs(a - b) <= rel * fabs(b) + abs_;
}

int main(void)
{
    double *tab = malloc(sizeof(double) * OB_BOYS_M * OB_BOYS_NX);
    ob_boys_table(tab);
    /* F_m(0) = 1/(2m+1), to the node error of the reference's quadrature
     * (F_0 is the closed form; F_1 carries the largest error, ~7e-6) */
    for (int m = 0; m < OB_BOYS_M; m++) {
        const double tol = m == 0 ? 0.0 : (m == 1 ? 1e-5 : (m < 6 ? 1e-8 : 1e-12));
        CHECK(close_rel(tab[m * OB_BOYS_NX], 1.0 / (2 * m + 1), tol, 0.0), "F_%d(0) = %.17g", m, tab[m * OB_BOYS_NX]);
    }
    for (int m = 0; m <= 8; m++)
        for (double x = 0.0; x < 130.0; x += 0.37) {
            const double t = ob_boys_taylor(tab, m, x), e = ob_boys_exact(m, x);
            CHECK(isfinite(t) && close_rel(t, e, 1e-5, 1e-12), "Boys m=%d x=%g: table %.17g accurate %.17g", m, x, t, e);
        }
    /* the positive half of the 256 Gauss-Legendre nodes: the weights add to
     * 1 and integrate x^2 to 1/3 and x^254 to 1/255 over [0, 1] */
    double glx[128], glw[128];
    ob_gl256(glx, glw);
    double w0 = 0.0, w2 = 0.0, w254 = 0.0;
    for (int i = 0; i < 128; i++) {
        w0 += glw[i];
        w2 += glw[i] * glx[i] * glx[i];
        w254 += glw[i] * pow(glx[i], 254);
    }
    CHECK(close_rel(w0, 1.0, 1e-14, 0.0), "GL-256 weights sum to %.17g", w0);
    CHECK(close_rel(w2, 1.0 / 3.0, 1e-14, 0.0), "GL-256 x^2 moment %.17g", w2);
    CHECK(close_rel(w254, 1.0 / 255.0, 1e-12, 0.0), "GL-256 x^254 moment %.17g", w254);

    ob_basis bs;
    make_basis(&bs);
    const long n = NBF, flat = n * (n + 1) / 2, len = ob_flat_len(n);

    /* 8-fold symmetry of contracted ERIs (accurate Boys, no screen) */
    for (int i = 0; i < n; i += 1)
        for (int j = 0; j < n; j += 2)
            for (int k = 1; k < n; k += 3)
                for (int l = 0; l < n; l += 2) {
                    const double v = ob_eri(&bs, tab, 1, 0, i, j, k, l, NULL);
                    const int perm[7][4] = {{j, i, k, l}, {i, j, l, k}, {j, i, l, k}, {k, l, i, j},
                                            {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
                    for (int p = 0; p < 7; p++) {
                        const double w = ob_eri(&bs, tab, 1, 0, perm[p][0], perm[p][1], perm[p][2], perm[p][3], NULL);
                        CHECK(close_rel(w, v, 1e-12, 1e-15), "(%d%d|%d%d) permutation %d: %.17g vs %.17g", i, j, k, l, p, w, v);
                    }
                }

    /* flat layout */
    for (long m = 1; m <= 8; m++) {
        long pos = 0;
        for (long i = 0; i < m; i++)
            for (long j = i; j < m; j++)
                for (long k = i; k < m; k++)
                    for (long l = k; l < m; l++) {
                        long in[4] = {i, j, k, l}, out[4];
                        ob_sort_keys(in, out);
                        if (out[0] != i || out[1] != j || out[2] != k || out[3] != l) continue;
                        CHECK(ob_flat_rank(m, i, j, k, l) == pos, "flat_rank n=%ld (%ld%ld|%ld%ld)", m, i, j, k, l);
                        CHECK(ob_four_index_ref(m, i, j, k, l) == pos, "fourIndex2Flat n=%ld (%ld%ld|%ld%ld)", m, i, j, k, l);
                        pos++;
                    }
        CHECK(pos == ob_flat_len(m), "flat length n=%ld", m);
    }

    /* flat arrays in both modes; G against the dense contraction */
    double *eri = malloc(sizeof(double) * len), *eri2 = malloc(sizeof(double) * len);
    long kept = 0, kept2 = 0;
    ob_eri_flat(&bs, tab, 0, 1, eri, &kept);
    ob_eri_flat(&bs, tab, 1, 0, eri2, &kept2);
    CHECK(kept > 0 && kept <= kept2, "screen kept %ld of %ld primitive quartets", kept, kept2);
    for (long t = 0; t < len; t++) CHECK(isfinite(eri[t]) && isfinite(eri2[t]), "flat entry %ld not finite", t);
    for (long i = 0; i < n; i++)
        for (long j = 0; j < n; j++)
            for (long k = 0; k < n; k++)
                for (long l = 0; l < n; l++) {
                    /* sortKeys gives the canonical quartet of any index order */
                    long in[4] = {i, j, k, l}, s[4];
                    ob_sort_keys(in, s);
                    const long r = ob_flat_rank(n, s[0], s[1], s[2], s[3]);
                    CHECK(ob_flat_get(n, eri2, i, j, k, l) == eri2[r], "flat_get (%ld%ld|%ld%ld)", i, j, k, l);
                }
    double *D = malloc(sizeof(double) * flat), *G = malloc(sizeof(double) * flat);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    for (long t = 0; t < flat; t++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        D[t] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
    }
    ob_gmatrix(n, eri2, D, G);
    for (long a = 0; a < n; a++)
        for (long b = a; b < n; b++) {
            double g = 0.0;
            for (long k = 0; k < n; k++)
                for (long l = 0; l < n; l++) {
                    const double dlk = D[ob_index_dim2_to_flat(n, l, k)];
                    g += dlk * (ob_flat_get(n, eri2, a, b, k, l) - 0.5 * ob_flat_get(n, eri2, a, l, k, b));
                }
            const double got = G[ob_index_dim2_to_flat(n, a, b)];
            CHECK(close_rel(got, g, 1e-12, 1e-13), "G[%ld,%ld] = %.17g, dense %.17g", a, b, got, g);
        }

    /* one-electron integrals and the quartet-list entry point */
    double *S = malloc(sizeof(double) * flat), *H = malloc(sizeof(double) * flat);
    const double atoms[9] = {0.0, 0.0, 0.1, 1.43, 0.0, -0.9, -1.43, 0.2, -0.9}, Z[3] = {8.0, 1.0, 1.0};
    ob_overlap(&bs, S);
    ob_hcore(&bs, tab, 0, 3, atoms, Z, H);
    for (long i = 0; i < n; i++) CHECK(S[ob_index_dim2_to_flat(n, i, i)] == 1.0, "S_%ld%ld != 1", i, i);
    for (long t = 0; t < flat; t++) CHECK(isfinite(S[t]) && isfinite(H[t]), "S/H entry %ld not finite", t);
    const int q[8] = {0, 1, 2, 3, 11, 10, 4, 9};
    double out[2];
    ob_eri_list(&bs, tab, 0, 1, 2, q, out);
    CHECK(out[0] == ob_eri(&bs, tab, 0, 1, 0, 1, 2, 3, NULL), "ob_eri_list entry 0");
    CHECK(out[1] == ob_eri(&bs, tab, 0, 1, 11, 10, 4, 9, NULL), "ob_eri_list entry 1");

    free(tab);
    free(eri);
    free(eri2);
    free(D);
    free(G);
    free(S);
    free(H);
    free(tls_rs);
    tls_rs = NULL;
    if (fails) {
        fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    printf("oracle sanitizer run ok: %ld flat entries, %ld of %ld primitive quartets kept\n", len, kept, kept2);
    return 0;
}
