// One-electron integrals S, H = T + sum_C (-Z_C) V_C over contracted Cartesian
// functions, following Integrals/IntegralsEvaluation.hs:169-280,385-414:
// overlap by 1-D Obara-Saika, kinetic via the OS overlap table, nuclear
// attraction by McMurchie-Davidson with the Boys function of the context.
// One thread per (i <= j) function pair.
#include <hip/hip_runtime.h>
#include <math.h>

#include "hfk_internal.hpp"

namespace {

constexpr int LM = 5;   // per-axis exponent bound (l <= 2, kinetic needs l+2)

// obaraSaika table s[m][n], m <= la, n <= lb (IntegralsEvaluation.hs:209-221)
__device__ void os_table(double gamma, double s00, double pax, double pbx, int la, int lb, double s[LM][LM])
{
    const double c = 1.0 / (2.0 * gamma);
    for (int m = 0; m <= la; ++m)
        for (int n = 0; n <= lb; ++n) {
            double v;
            if (m == 0 && n == 0) v = s00;
            else if (m == 0) v = pbx * s[0][n - 1] + c * (double)(n - 1) * (n >= 2 ? s[0][n - 2] : 0.0);
            else if (n == 0) v = pax * s[m - 1][0] + c * (double)(m - 1) * (m >= 2 ? s[m - 2][0] : 0.0);
            else
                v = pax * s[m - 1][n] +
                    c * ((double)(m - 1) * (m >= 2 ? s[m - 2][n] : 0.0) + (double)n * s[m - 1][n - 1]);
            s[m][n] = v;
        }
}

__device__ void pair_geom(const double* r1, double e1, const double* r2, double e2, double& gamma, double s00[3],
                          double pa[3], double pb[3])
{
    gamma = e1 + e2;
    const double mu = e1 * e2 / gamma, cte = sqrt(HFK_PI / gamma);
    for (int x = 0; x < 3; ++x) {
        const double p = (e1 * r1[x] + e2 * r2[x]) / (e1 + e2);
        const double d = r1[x] - r2[x];
        s00[x] = cte * exp(-(mu * (d * d)));
        pa[x] = p - r1[x];
        pb[x] = p - r2[x];
    }
}

__device__ void flat_ij(long long idx, int n, int& i, int& j)
{
    const double nn = (double)n + 0.5;
    long long ii = (long long)(nn - sqrt(nn * nn - 2.0 * (double)idx));
    if (ii < 0) ii = 0;
    while (ii > 0 && ii * n - ii * (ii - 1) / 2 > idx) --ii;
    while ((ii + 1) * n - (ii + 1) * ii / 2 <= idx) ++ii;
    i = (int)ii;
    j = (int)(ii + (idx - (ii * n - ii * (ii - 1) / 2)));
}

__global__ void k_overlap(OneEArgs a, double* S)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        int i, j;
        flat_ij(idx, a.n, i, j);
        if (i == j) { S[idx] = 1.0; continue; }   // sijContracted: S_ii := 1
        const double* r1 = a.center + 3 * i;
        const double* r2 = a.center + 3 * j;
        const int* l1 = a.lxyz + 3 * i;
        const int* l2 = a.lxyz + 3 * j;
        double sum = 0.0;
        for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
            for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2) {
                double gamma, s00[3], pa[3], pb[3];
                pair_geom(r1, a.expo[g1], r2, a.expo[g2], gamma, s00, pa, pb);
                double pr = 1.0;
                for (int x = 0; x < 3; ++x) {
                    double s[LM][LM];
                    os_table(gamma, s00[x], pa[x], pb[x], l1[x], l2[x], s);
                    pr *= s[l1[x]][l2[x]];
                }
                sum += (a.coef[g1] * a.coef[g2]) * pr;
            }
        S[idx] = sum;
    }
}

// Hermite coefficients E[i][j][t] per axis, seed e0 (IntegralsEvaluation.hs:457-477)
__device__ void herm_axis1(double gamma, double xpa, double xpb, double e0, int li, int lj, double E[3][3][5])
{
    for (int i = 0; i <= li; ++i)
        for (int j = 0; j <= lj; ++j)
            for (int t = 0; t <= 4; ++t) {
                double v;
                if (i == 0 && j == 0 && t == 0) v = e0;
                else if (t > i + j) v = 0.0;
                else if (t == 0 && i >= 1) v = xpa * E[i - 1][j][0] + E[i - 1][j][1];
                else if (t == 0 && j >= 1) v = xpb * E[i][j - 1][0] + E[i][j - 1][1];
                else {
                    const double av = (i >= 1) ? E[i - 1][j][t - 1] : 0.0;
                    const double bv = (j >= 1) ? E[i][j - 1][t - 1] : 0.0;
                    v = (1.0 / (2.0 * gamma * (double)t)) * ((double)i * av + (double)j * bv);
                }
                E[i][j][t] = v;
            }
}

template <int BOYS>
__device__ double nuclear_prim(const double* ra, const int* la, double c1, double e1, const double* rb,
                               const int* lb, double c2, double e2, const double* rc, const BoysTables& bt)
{
    const double gamma = e1 + e2, nu = e1 * e2 / gamma;
    double rp[3], rpc[3];
    double E[3][3][3][5];
    for (int x = 0; x < 3; ++x) {
        rp[x] = (e1 * ra[x] + e2 * rb[x]) / (e1 + e2);
        rpc[x] = rp[x] - rc[x];
        const double d = ra[x] - rb[x];
        herm_axis1(gamma, rp[x] - ra[x], rp[x] - rb[x], exp(-(nu * (d * d))), la[x], lb[x], E[x]);
    }
    const int L = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2];
    const double T = gamma * ((rpc[0] * rpc[0] + rpc[1] * rpc[1]) + rpc[2] * rpc[2]);
    double F[5];
    boys_eval<double, 4, BOYS>(T, bt, F);
    // R^n_{tuv}, t+u+v <= L-n, layered from n = L down to 0 (IntegralsEvaluation.hs:430-449)
    double Rn[5][5][5], Rm[5][5][5];
    double m2g = 1.0;
    for (int n = 0; n < L; ++n) m2g *= -2.0 * gamma;
    for (int n = L; n >= 0; --n) {
        for (int t = 0; t <= L - n; ++t)
            for (int u = 0; u <= L - n - t; ++u)
                for (int v = 0; v <= L - n - t - u; ++v) {
                    double r;
                    if (t >= 1) r = (t >= 2 ? (double)(t - 1) * Rm[t - 2][u][v] : 0.0) + rpc[0] * Rm[t - 1][u][v];
                    else if (u >= 1) r = (u >= 2 ? (double)(u - 1) * Rm[t][u - 2][v] : 0.0) + rpc[1] * Rm[t][u - 1][v];
                    else if (v >= 1) r = (v >= 2 ? (double)(v - 1) * Rm[t][u][v - 2] : 0.0) + rpc[2] * Rm[t][u][v - 1];
                    else r = m2g * F[n];
                    Rn[t][u][v] = r;
                }
        for (int t = 0; t <= L - n; ++t)
            for (int u = 0; u <= L - n - t; ++u)
                for (int v = 0; v <= L - n - t - u; ++v) Rm[t][u][v] = Rn[t][u][v];
        if (n > 0) m2g /= (-2.0 * gamma);
    }
    double s = 0.0;
    for (int t = 0; t <= la[0] + lb[0]; ++t)
        for (int u = 0; u <= la[1] + lb[1]; ++u)
            for (int v = 0; v <= la[2] + lb[2]; ++v)
                s += ((1.0 * E[0][la[0]][lb[0]][t]) * E[1][la[1]][lb[1]][u]) * E[2][la[2]][lb[2]][v] * Rm[t][u][v];
    return c1 * c2 * 2.0 * (HFK_PI / gamma) * s;
}

template <int BOYS>
__global__ void k_hcore(OneEArgs a, double* H)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        int i, j;
        flat_ij(idx, a.n, i, j);
        const double* r1 = a.center + 3 * i;
        const double* r2 = a.center + 3 * j;
        const int* l1 = a.lxyz + 3 * i;
        const int* l2 = a.lxyz + 3 * j;
        // kinetic: tab (IntegralsEvaluation.hs:263-280)
        double tsum = 0.0;
        for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
            for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2) {
                const double e2 = a.expo[g2];
                double gamma, s00[3], pa[3], pb[3];
                pair_geom(r1, a.expo[g1], r2, e2, gamma, s00, pa, pb);
                double sx[3], tx[3];
                for (int x = 0; x < 3; ++x) {
                    double s[LM][LM];
                    os_table(gamma, s00[x], pa[x], pb[x], l1[x], l2[x] + 2, s);
                    const double L2 = (double)l2[x];
                    sx[x] = s[l1[x]][l2[x]];
                    tx[x] = -2.0 * (e2 * e2) * s[l1[x]][l2[x] + 2] + e2 * (2.0 * L2 + 1.0) * sx[x] -
                            0.5 * L2 * (L2 - 1.0) * (l2[x] >= 2 ? s[l1[x]][l2[x] - 2] : 0.0);
                }
                const double t1 = tx[0] * sx[1] * sx[2];
                const double t2 = sx[0] * tx[1] * sx[2];
                const double t3 = sx[0] * sx[1] * tx[2];
                tsum += a.coef[g1] * a.coef[g2] * ((t1 + t2) + t3);
            }
        double vsum = 0.0;
        for (int c = 0; c < a.natom; ++c) {
            double v = 0.0;
            for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
                for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2)
                    v += nuclear_prim<BOYS>(r1, l1, a.coef[g1], a.expo[g1], r2, l2, a.coef[g2], a.expo[g2],
                                            a.R + 3 * c, a.bt);
            vsum += (-a.Z[c]) * v;
        }
        H[idx] = tsum + vsum;
    }
}


// ---------------------------------------------------------------------
// One-electron part of the nuclear gradient.  For a function pair (i, j)
// on centres A, B the derivative of a primitive with respect to its own
// centre is d/dA_k g_l = 2 a g_{l+1_k} - l_k g_{l-1_k}, so every
// derivative integral is a combination of ordinary integrals with shifted
// exponents.  S and T depend on A - B only (dB = -dA); the nuclear
// attraction of nucleus C also moves with C (dC = -(dA + dB)).
// ---------------------------------------------------------------------

// primitive overlap and kinetic integrals for explicit exponents (l <= 3 on
// the first function, <= 2 on the second)
__device__ void st_prim(const double* r1, const int* l1, double e1, const double* r2, const int* l2, double e2,
                        double& sval, double& tval)
{
    double gamma, s00[3], pa[3], pb[3];
    pair_geom(r1, e1, r2, e2, gamma, s00, pa, pb);
    double sx[3], tx[3];
    for (int x = 0; x < 3; ++x) {
        double s[LM][LM];
        os_table(gamma, s00[x], pa[x], pb[x], l1[x], l2[x] + 2, s);
        const double L2 = (double)l2[x];
        sx[x] = s[l1[x]][l2[x]];
        tx[x] = -2.0 * (e2 * e2) * s[l1[x]][l2[x] + 2] + e2 * (2.0 * L2 + 1.0) * sx[x] -
                0.5 * L2 * (L2 - 1.0) * (l2[x] >= 2 ? s[l1[x]][l2[x] - 2] : 0.0);
    }
    sval = sx[0] * sx[1] * sx[2];
    tval = (tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2]) + sx[0] * sx[1] * tx[2];
}

// Hermite coefficients with one extra unit of angular momentum per centre
__device__ void herm_axis_g(double gamma, double xpa, double xpb, double e0, int li, int lj, double E[4][4][7])
{
    for (int i = 0; i <= li; ++i)
        for (int j = 0; j <= lj; ++j)
            for (int t = 0; t <= 6; ++t) {
                double v;
                if (i == 0 && j == 0 && t == 0) v = e0;
                else if (t > i + j) v = 0.0;
                else if (t == 0 && i >= 1) v = xpa * E[i - 1][j][0] + E[i - 1][j][1];
                else if (t == 0 && j >= 1) v = xpb * E[i][j - 1][0] + E[i][j - 1][1];
                else {
                    const double av = (i >= 1) ? E[i - 1][j][t - 1] : 0.0;
                    const double bv = (j >= 1) ? E[i][j - 1][t - 1] : 0.0;
                    v = (1.0 / (2.0 * gamma * (double)t)) * ((double)i * av + (double)j * bv);
                }
                E[i][j][t] = v;
            }
}

// d/dA and d/dB of the primitive nuclear-attraction integral <a|1/r_C|b>
// (without -Z_C), McMurchie-Davidson with R_tuv to L+1.
template <int BOYS>
__device__ void nuclear_prim_grad(const double* ra, const int* la, double e1, const double* rb, const int* lb,
                                  double e2, const double* rc, const BoysTables& bt, double dA[3], double dB[3])
{
    const double gamma = e1 + e2, nu = e1 * e2 / gamma;
    double rpc[3];
    double E[3][4][4][7];
    for (int x = 0; x < 3; ++x) {
        const double rp = (e1 * ra[x] + e2 * rb[x]) / (e1 + e2);
        rpc[x] = rp - rc[x];
        const double d = ra[x] - rb[x];
        herm_axis_g(gamma, rp - ra[x], rp - rb[x], exp(-(nu * (d * d))), la[x] + 1, lb[x] + 1, E[x]);
    }
    const int L = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2] + 1;
    const double T = gamma * ((rpc[0] * rpc[0] + rpc[1] * rpc[1]) + rpc[2] * rpc[2]);
    double F[7];
    boys_eval<double, 6, BOYS>(T, bt, F);
    double Rn[7][7][7], Rm[7][7][7];
    double m2g = 1.0;
    for (int n = 0; n < L; ++n) m2g *= -2.0 * gamma;
    for (int n = L; n >= 0; --n) {
        for (int t = 0; t <= L - n; ++t)
            for (int u = 0; u <= L - n - t; ++u)
                for (int v = 0; v <= L - n - t - u; ++v) {
                    double r;
                    if (t >= 1) r = (t >= 2 ? (double)(t - 1) * Rm[t - 2][u][v] : 0.0) + rpc[0] * Rm[t - 1][u][v];
                    else if (u >= 1) r = (u >= 2 ? (double)(u - 1) * Rm[t][u - 2][v] : 0.0) + rpc[1] * Rm[t][u - 1][v];
                    else if (v >= 1) r = (v >= 2 ? (double)(v - 1) * Rm[t][u][v - 2] : 0.0) + rpc[2] * Rm[t][u][v - 1];
                    else r = m2g * F[n];
                    Rn[t][u][v] = r;
                }
        for (int t = 0; t <= L - n; ++t)
            for (int u = 0; u <= L - n - t; ++u)
                for (int v = 0; v <= L - n - t - u; ++v) Rm[t][u][v] = Rn[t][u][v];
        if (n > 0) m2g /= (-2.0 * gamma);
    }
    const double pref = 2.0 * (HFK_PI / gamma);
    for (int k = 0; k < 3; ++k) {
        for (int side = 0; side < 2; ++side) {
            // derivative coefficients along axis k, plain ones on the others
            double Dk[7];
            const int i = la[k], j = lb[k];
            const int tmax = i + j + 1;
            for (int t = 0; t <= tmax; ++t) {
                double v;
                if (side == 0) v = 2.0 * e1 * E[k][i + 1][j][t] - (i >= 1 ? (double)i * E[k][i - 1][j][t] : 0.0);
                else v = 2.0 * e2 * E[k][i][j + 1][t] - (j >= 1 ? (double)j * E[k][i][j - 1][t] : 0.0);
                Dk[t] = v;
            }
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            double s = 0.0;
            for (int t = 0; t <= tmax; ++t)
                for (int u = 0; u <= la[k1] + lb[k1]; ++u)
                    for (int v = 0; v <= la[k2] + lb[k2]; ++v) {
                        int idx[3];
                        idx[k] = t;
                        idx[k1] = u;
                        idx[k2] = v;
                        s += Dk[t] * E[k1][la[k1]][lb[k1]][u] * E[k2][la[k2]][lb[k2]][v] * Rm[idx[0]][idx[1]][idx[2]];
                    }
            if (side == 0) dA[k] = pref * s;
            else dB[k] = pref * s;
        }
    }
}

__device__ __forceinline__ void grad_add(double* acc, int i, double v)
{
    if (v != 0.0) atomic_add_f64(&acc[i], v);
}

template <int BOYS>
__global__ void __launch_bounds__(128) k_grad_onee(OneEArgs a, const int* atom_of_bf, const double* Dflat,
                                                   const double* Wflat, double* grad)
{
    extern __shared__ double gacc[];   // [natom][3]
    for (int t = threadIdx.x; t < 3 * a.natom; t += blockDim.x) gacc[t] = 0.0;
    __syncthreads();
    const long long m = (long long)a.n * (a.n + 1) / 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        int i, j;
        flat_ij(idx, a.n, i, j);
        const double w2 = (i == j) ? 1.0 : 2.0;   // (i,j) and (j,i)
        const double wd = w2 * Dflat[idx], ww = w2 * Wflat[idx];
        const double* r1 = a.center + 3 * i;
        const double* r2 = a.center + 3 * j;
        const int* l1 = a.lxyz + 3 * i;
        const int* l2 = a.lxyz + 3 * j;
        const int ai = atom_of_bf[i], aj = atom_of_bf[j];
        if (ai != aj) {
            // kinetic and overlap (Pulay): dA only, dB = -dA
            double dA[3] = {0.0, 0.0, 0.0};
            for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
                for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2) {
                    const double e1 = a.expo[g1], e2 = a.expo[g2], cc = a.coef[g1] * a.coef[g2];
                    for (int k = 0; k < 3; ++k) {
                        int lp[3] = {l1[0], l1[1], l1[2]};
                        lp[k] += 1;
                        double sp, tp, sm = 0.0, tm = 0.0;
                        st_prim(r1, lp, e1, r2, l2, e2, sp, tp);
                        if (l1[k] >= 1) {
                            lp[k] -= 2;
                            st_prim(r1, lp, e1, r2, l2, e2, sm, tm);
                        }
                        const double ds = 2.0 * e1 * sp - (double)l1[k] * sm;
                        const double dt = 2.0 * e1 * tp - (double)l1[k] * tm;
                        dA[k] += cc * (wd * dt - ww * ds);
                    }
                }
            for (int k = 0; k < 3; ++k) {
                grad_add(gacc, 3 * ai + k, dA[k]);
                grad_add(gacc, 3 * aj + k, -dA[k]);
            }
        }
        // nuclear attraction: basis-centre and Hellmann-Feynman terms
        for (int c = 0; c < a.natom; ++c) {
            double dA[3] = {0.0, 0.0, 0.0}, dB[3] = {0.0, 0.0, 0.0};
            for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
                for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2) {
                    double pa[3], pb[3];
                    nuclear_prim_grad<BOYS>(r1, l1, a.expo[g1], r2, l2, a.expo[g2], a.R + 3 * c, a.bt, pa, pb);
                    const double cc = a.coef[g1] * a.coef[g2];
                    for (int k = 0; k < 3; ++k) {
                        dA[k] += cc * pa[k];
                        dB[k] += cc * pb[k];
                    }
                }
            const double zw = -a.Z[c] * wd;
            for (int k = 0; k < 3; ++k) {
                grad_add(gacc, 3 * ai + k, zw * dA[k]);
                grad_add(gacc, 3 * aj + k, zw * dB[k]);
                grad_add(gacc, 3 * c + k, -zw * (dA[k] + dB[k]));
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * a.natom; t += blockDim.x)
        if (gacc[t] != 0.0) atomic_add_f64(&grad[t], gacc[t]);
}

unsigned grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

void hfk_launch_overlap(const OneEArgs& a, double* S_flat, hipStream_t s)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    hipLaunchKernelGGL(k_overlap, dim3(grid_for(m, 128)), dim3(128), 0, s, a, S_flat);
}

void hfk_launch_hcore(const OneEArgs& a, double* H_flat, hipStream_t s)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    if (a.boys == 0)
        hipLaunchKernelGGL(k_hcore<0>, dim3(grid_for(m, 64)), dim3(64), 0, s, a, H_flat);
    else
        hipLaunchKernelGGL(k_hcore<1>, dim3(grid_for(m, 64)), dim3(64), 0, s, a, H_flat);
}

void hfk_launch_grad_onee(const OneEArgs& a, const int* atom_of_bf, const double* D_flat, const double* W_flat,
                          double* grad, hipStream_t s)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    const size_t shm = sizeof(double) * 3 * (size_t)a.natom;
    long long g = (m + 127) / 128;
    if (g > 2048) g = 2048;   // bounded: each workgroup flushes natom*3 sums
    if (a.boys == 0)
        hipLaunchKernelGGL(k_grad_onee<0>, dim3((unsigned)g), dim3(128), shm, s, a, atom_of_bf, D_flat, W_flat, grad);
    else
        hipLaunchKernelGGL(k_grad_onee<1>, dim3((unsigned)g), dim3(128), shm, s, a, atom_of_bf, D_flat, W_flat, grad);
}
