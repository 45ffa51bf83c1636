// One-electron integrals S, H = T + sum_C (-Z_C) V_C over contracted Cartesian
// functions, following Integrals/IntegralsEvaluation.hs:169-280,385-414:
// overlap by 1-D Obara-Saika, kinetic via the OS overlap table, nuclear
// attraction by McMurchie-Davidson with the Boys function of the context.
// One thread per (i <= j) function pair.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

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

__device__ __forceinline__ void grad_add(double* acc, int i, double v)
{
    if (v != 0.0) atomic_add_f64(&acc[i], v);
}

// Kinetic and overlap (Pulay) terms, one thread per function pair.
__global__ void __launch_bounds__(128) k_grad_onee_st(OneEArgs a, const int* atom_of_bf, const double* Dflat,
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
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * a.natom; t += blockDim.x)
        if (gacc[t] != 0.0) atomic_add_f64(&grad[t], gacc[t]);
}


// ---------------------------------------------------------------------
// Nuclear attraction V_ij = sum_C -Z_C <i|1/r_C|j> (McMurchie-Davidson,
// IntegralsEvaluation.hs:385-414) and its derivatives with respect to the
// basis centres and the nuclei.  Function pairs are grouped by
// l_i + l_j = LS so that R_tuv has a compile-time order and stays in
// registers; a thread takes one pair and one nucleus.  Per primitive pair
// the per-axis Hermite tables are built unrolled, the pair's own 1-D
// coefficients are picked by selects, and the contraction runs densely over
// t+u+v <= LR with zero padding.
// ---------------------------------------------------------------------

constexpr int oe_nherm(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
__host__ __device__ constexpr int oe_hidx(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

// E^{ij}_t, i, j <= IM, Helgaker's recursion with E^{00}_0 = 1
template <int IM>
__device__ __forceinline__ void oe_herm(double PA, double PB, double oo2p, double (&E)[IM + 1][IM + 1][2 * IM + 1])
{
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= IM; ++j)
#pragma unroll
            for (int t = 0; t <= 2 * IM; ++t) {
                double v = 0.0;
                if (i == 0 && j == 0) {
                    v = (t == 0) ? 1.0 : 0.0;
                } else if (t <= i + j) {
                    const int si = i > 0 ? i - 1 : 0, sj = i > 0 ? j : j - 1;
                    const double X = i > 0 ? PA : PB;
                    if (t >= 1) v += oo2p * E[si][sj][t - 1];
                    if (t <= si + sj) v += X * E[si][sj][t];
                    if (t + 1 <= si + sj) v += (double)(t + 1) * E[si][sj][t + 1];
                }
                E[i][j][t] = v;
            }
}

// the pair's 1-D coefficients along one axis (i0 on A, j0 on B): E^{i0 j0}_t
// and, with DERIV, 2a E^{i0+1,j0} - i0 E^{i0-1,j0} and 2b E^{i0,j0+1} - j0 E^{i0,j0-1}
template <int IM, int TN, bool DERIV>
__device__ __forceinline__ void oe_axis(const double (&E)[IM + 1][IM + 1][2 * IM + 1], int i0, int j0, double a2,
                                        double b2, double* e, double* da, double* db)
{
#pragma unroll
    for (int t = 0; t < TN; ++t) {
        double v = 0.0, pi = 0.0, mi = 0.0, pj = 0.0, mj = 0.0;
        if (t <= 2 * IM) {
#pragma unroll
            for (int ii = 0; ii <= IM; ++ii)
#pragma unroll
                for (int jj = 0; jj <= IM; ++jj) {
                    const double x = E[ii][jj][t];
                    v = (ii == i0 && jj == j0) ? x : v;
                    if constexpr (DERIV) {
                        pi = (ii == i0 + 1 && jj == j0) ? x : pi;
                        mi = (ii == i0 - 1 && jj == j0) ? x : mi;
                        pj = (ii == i0 && jj == j0 + 1) ? x : pj;
                        mj = (ii == i0 && jj == j0 - 1) ? x : mj;
                    }
                }
        }
        e[t] = v;
        if constexpr (DERIV) {
            da[t] = a2 * pi - (double)i0 * mi;
            db[t] = b2 * pj - (double)j0 * mj;
        }
    }
}

// R_tuv, t+u+v <= LR, from F[n] = pref (-2 gamma)^n F_n
template <int LR>
__device__ __forceinline__ void oe_R(const double* F, double X, double Y, double Z, double* R)
{
    double cur[oe_nherm(LR)];
    cur[0] = F[LR];
#pragma unroll
    for (int n = LR - 1; n >= 0; --n) {
        double nxt[oe_nherm(LR)];
        const int top = LR - n;
#pragma unroll
        for (int t = 0; t <= LR; ++t)
#pragma unroll
            for (int u = 0; u <= LR - t; ++u)
#pragma unroll
                for (int v = 0; v <= LR - t - u; ++v) {
                    if (t + u + v > top) continue;
                    double val;
                    if (t > 0) val = (t >= 2 ? (double)(t - 1) * cur[oe_hidx(t - 2, u, v)] : 0.0) + X * cur[oe_hidx(t - 1, u, v)];
                    else if (u > 0) val = (u >= 2 ? (double)(u - 1) * cur[oe_hidx(t, u - 2, v)] : 0.0) + Y * cur[oe_hidx(t, u - 1, v)];
                    else if (v > 0) val = (v >= 2 ? (double)(v - 1) * cur[oe_hidx(t, u, v - 2)] : 0.0) + Z * cur[oe_hidx(t, u, v - 1)];
                    else val = F[n];
                    nxt[oe_hidx(t, u, v)] = val;
                }
#pragma unroll
        for (int e = 0; e < oe_nherm(LR); ++e)
            if (e < oe_nherm(top)) cur[e] = nxt[e];
    }
#pragma unroll
    for (int e = 0; e < oe_nherm(LR); ++e) R[e] = cur[e];
}

// !DERIV: H[idx] += sum_C -Z_C V^C_ij, one thread per function pair summing
// the nuclei in order (a fixed-order sum: hcore is bitwise reproducible).
// DERIV: one thread per (function pair, nucleus); sum D_ij dV_ij/dR into grad.
template <int LS, int BOYS, bool DERIV>
__global__ void __launch_bounds__(128) k_onee_v(OneEArgs a, const int* plist, int np, int ng, const int* atom_of_bf,
                                                const double* Dflat, double* H, double* grad)
{
    constexpr int LR = LS + (DERIV ? 1 : 0);
    constexpr int TN = LS + 2;
    constexpr int IM = (LS < 2 ? LS : 2) + (DERIV ? 1 : 0);   // per-axis exponent bound of either function
    extern __shared__ double gacc[];   // [natom][3] (DERIV)
    if constexpr (DERIV) {
        for (int t = threadIdx.x; t < 3 * a.natom; t += blockDim.x) gacc[t] = 0.0;
        __syncthreads();
    }
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long p = tid / ng;
    const int c0 = DERIV ? (int)(tid - p * ng) : 0;   // the nucleus of this thread (DERIV), else all
    const int c1 = DERIV ? c0 + 1 : a.natom;
    double hsum = 0.0;
    if (p < np) for (int c = c0; c < c1; ++c) {
        const long long idx = plist[p];
        int i, j;
        flat_ij(idx, a.n, i, j);
        const double* r1 = a.center + 3 * i;
        const double* r2 = a.center + 3 * j;
        const int* l1 = a.lxyz + 3 * i;
        const int* l2 = a.lxyz + 3 * j;
        const double Cx = a.R[3 * c], Cy = a.R[3 * c + 1], Cz = a.R[3 * c + 2];
        double hv = 0.0;
        double dA[3] = {0.0, 0.0, 0.0}, dB[3] = {0.0, 0.0, 0.0};
        // derivatives: primitive pairs with exp(-mu |AB|^2) < 1e-20 (mu |AB|^2 > 46)
        // are skipped; the potential itself (hcore) keeps every pair
        const double dab2 = (r1[0] - r2[0]) * (r1[0] - r2[0]) + (r1[1] - r2[1]) * (r1[1] - r2[1]) +
                            (r1[2] - r2[2]) * (r1[2] - r2[2]);
        for (int g1 = a.prim_off[i]; g1 < a.prim_off[i + 1]; ++g1)
            for (int g2 = a.prim_off[j]; g2 < a.prim_off[j + 1]; ++g2) {
                const double e1 = a.expo[g1], e2 = a.expo[g2];
                const double gamma = e1 + e2, rg = 1.0 / gamma, mu = e1 * e2 * rg;
                if constexpr (DERIV) {
                    if (mu * dab2 > 46.0) continue;
                }
                double P[3], ex[3][TN], dax[3][TN], dbx[3][TN], ab2 = 0.0;
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    P[x] = (e1 * r1[x] + e2 * r2[x]) * rg;
                    const double d = r1[x] - r2[x];
                    ab2 += d * d;
                    double E[IM + 1][IM + 1][2 * IM + 1];
                    oe_herm<IM>(P[x] - r1[x], P[x] - r2[x], 0.5 * rg, E);
                    oe_axis<IM, TN, DERIV>(E, l1[x], l2[x], 2.0 * e1, 2.0 * e2, ex[x], dax[x], dbx[x]);
                }
                const double X = P[0] - Cx, Y = P[1] - Cy, Z = P[2] - Cz;
                double F[LR + 1];
                boys_eval<double, LR, BOYS>(gamma * ((X * X + Y * Y) + Z * Z), a.bt, F);
                double w = a.coef[g1] * a.coef[g2] * 2.0 * HFK_PI * rg * exp(-mu * ab2);
#pragma unroll
                for (int n = 0; n <= LR; ++n) {
                    F[n] *= w;
                    w *= -2.0 * gamma;
                }
                double R[oe_nherm(LR)];
                oe_R<LR>(F, X, Y, Z, R);
                if constexpr (!DERIV) {
#pragma unroll
                    for (int t = 0; t <= LS; ++t)
#pragma unroll
                        for (int u = 0; u <= LS - t; ++u)
#pragma unroll
                            for (int q = 0; q <= LS - t - u; ++q)
                                hv += ex[0][t] * ex[1][u] * ex[2][q] * R[oe_hidx(t, u, q)];
                } else {
#pragma unroll
                    for (int t = 0; t <= LR; ++t)
#pragma unroll
                        for (int u = 0; u <= LR - t; ++u)
#pragma unroll
                            for (int q = 0; q <= LR - t - u; ++q) {
                                const double r = R[oe_hidx(t, u, q)];
                                dA[0] += dax[0][t] * ex[1][u] * ex[2][q] * r;
                                dA[1] += ex[0][t] * dax[1][u] * ex[2][q] * r;
                                dA[2] += ex[0][t] * ex[1][u] * dax[2][q] * r;
                                dB[0] += dbx[0][t] * ex[1][u] * ex[2][q] * r;
                                dB[1] += ex[0][t] * dbx[1][u] * ex[2][q] * r;
                                dB[2] += ex[0][t] * ex[1][u] * dbx[2][q] * r;
                            }
                }
            }
        const double zc = -a.Z[c];
        if constexpr (!DERIV) {
            hsum += zc * hv;
            if (c == c1 - 1) H[idx] += hsum;
        } else {
            const double s = zc * ((i == j) ? 1.0 : 2.0) * Dflat[idx];
            const int ai = atom_of_bf[i], aj = atom_of_bf[j];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                atomic_add_f64(&gacc[3 * ai + x], s * dA[x]);
                atomic_add_f64(&gacc[3 * aj + x], s * dB[x]);
                atomic_add_f64(&gacc[3 * c + x], -s * (dA[x] + dB[x]));
            }
        }
    }
    if constexpr (DERIV) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * a.natom; t += blockDim.x)
            if (gacc[t] != 0.0) atomic_add_f64(&grad[t], gacc[t]);
    }
}

// Kinetic energy T_ij (IntegralsEvaluation.hs:263-280), written as H[idx] = T_ij.
__global__ void k_kinetic(OneEArgs a, double* H)
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
        H[idx] = tsum;
    }
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

template <bool DERIV>
static void launch_onee_v(const OneEArgs& a, const int* atom_of_bf, const double* D_flat, double* H, double* grad,
                          hipStream_t s)
{
    const int ng = DERIV ? a.natom : 1;   // one thread per (function pair, nucleus) / per function pair
    const size_t shm = DERIV ? sizeof(double) * 3 * (size_t)a.natom : 0;
    for (int ls = 0; ls <= 4; ++ls) {
        const int np = a.vpair_off[ls + 1] - a.vpair_off[ls];
        if (np <= 0) continue;
        const int* pl = a.vpair + a.vpair_off[ls];
        const long long nt = (long long)np * ng;
        const unsigned grid = (unsigned)((nt + 127) / 128);
#define HFK_V(LS)                                                                                              \
        if (ls == LS) {                                                                                       \
            if (a.boys == 0)                                                                                  \
                hipLaunchKernelGGL((k_onee_v<LS, 0, DERIV>), dim3(grid), dim3(128), shm, s, a, pl, np, ng, atom_of_bf, \
                                   D_flat, H, grad);                                                          \
            else                                                                                              \
                hipLaunchKernelGGL((k_onee_v<LS, 1, DERIV>), dim3(grid), dim3(128), shm, s, a, pl, np, ng, atom_of_bf, \
                                   D_flat, H, grad);                                                          \
        }
        HFK_V(0) HFK_V(1) HFK_V(2) HFK_V(3) HFK_V(4)
#undef HFK_V
    }
}

void hfk_launch_hcore(const OneEArgs& a, double* H_flat, hipStream_t s)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    hipLaunchKernelGGL(k_kinetic, dim3(grid_for(m, 64)), dim3(64), 0, s, a, H_flat);   // H = T
    launch_onee_v<false>(a, nullptr, nullptr, H_flat, nullptr, s);                      // H += V
}

void hfk_launch_grad_onee(const OneEArgs& a, const int* atom_of_bf, const double* D_flat, const double* W_flat,
                          double* grad, hipStream_t s)
{
    const long long m = (long long)a.n * (a.n + 1) / 2;
    const size_t shm = sizeof(double) * 3 * (size_t)a.natom;
    long long g = (m + 127) / 128;
    if (g > 2048) g = 2048;   // bounded: each workgroup flushes natom*3 sums
    hipLaunchKernelGGL(k_grad_onee_st, dim3((unsigned)g), dim3(128), shm, s, a, atom_of_bf, D_flat, W_flat, grad);
    launch_onee_v<true>(a, atom_of_bf, D_flat, nullptr, grad, s);
}
