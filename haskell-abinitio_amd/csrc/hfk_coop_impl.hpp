// Cooperative shell-quartet kernel for the classes with d shells, (ds|ss)
// through (dd|dd).  A class like (dd|dd) holds 1296 contracted integrals per
// quartet, too many for one lane's registers, so here a 64-lane wave
// evaluates quartets together:
//   * lane = (slot, bra Cartesian component ab); S = 64 / (na nb) slots take
//     S kets of the task at once (one ket per slot), all with the same bra;
//   * the bra's Hermite coefficient tables E^{ij}_t (per axis, per bra
//     primitive pair) and, per slot, the ket's tables and the Hermite
//     integrals R_tuv live in LDS; R is built level by level from the Boys
//     values (IntegralsEvaluation.hs:430-449 recursion) by the slot's lanes;
//   * each lane forms X^{ab}_{tau nu phi} = sum_tuv E^{ab}_tuv R_{t+tau,...}
//     and contracts it with every ket component:
//       (ab|cd) = sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} X^{ab}_{tau nu phi}
//     -- the McMurchie-Davidson product of the reference (twoElectronHermite,
//     mcMurchie2: IntegralsEvaluation.hs:340-376), summed over primitives
//     with the reference's per-component screen (:305-315).
// The same kernel digests into J/K (K rows of the bra shells staged in
// LDS), writes the flat calcIntegrals array, forms the Schwarz diagonal,
// or contracts derivative integrals with the two-particle density for the
// nuclear gradient (d/dA g_a = 2 alpha g_{a+1} - a g_{a-1}; D by
// translational invariance).  FP64 throughout.
#pragma once
#include "hfk_quartet_common.hpp"

namespace {

constexpr int nherm(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

// Position of (t,u,v) among all Hermite indices ordered by t+u+v, then by
// descending t, then by v; independent of the maximum order.
__host__ __device__ constexpr int hidx(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

__device__ __forceinline__ void hdecode(int e, int& t, int& u, int& v)
{
    int n = 0;
    while (nherm(n) <= e) ++n;
    const int r = e - nherm(n - 1);
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= r) ++m;
    v = r - m * (m + 1) / 2;
    u = m - v;
    t = n - m;
}

// exponent along `axis` of Cartesian component k of angular momentum l
// (x^l first, decreasing x, then decreasing y)
__host__ __device__ constexpr int cart_e(int l, int k, int axis)
{
    int c = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (c == k) return axis == 0 ? i : (axis == 1 ? j : l - i - j);
            ++c;
        }
    return 0;
}

// E^{ij}_t for i <= IM, j <= JM at [(i*(JM+1)+j)*(IM+JM+1)+t]; Helgaker's
// recursion with E^{00}_0 = 1 (the Gaussian product factor is in PP_K).
// Built in registers (fully unrolled), then stored.
template <int IM, int JM>
__device__ __forceinline__ void herm_table(double PA, double PB, double oo2p, double* E)
{
    constexpr int TS = IM + JM + 1, SZ = (IM + 1) * (JM + 1) * TS;
    double e[SZ];
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j)
#pragma unroll
            for (int t = 0; t < TS; ++t) {
                double v = 0.0;
                if (i == 0 && j == 0) {
                    v = (t == 0) ? 1.0 : 0.0;
                } else {
                    const int si = i > 0 ? i - 1 : i, sj = i > 0 ? j : j - 1;
                    const double X = i > 0 ? PA : PB;
                    const int smax = si + sj;
                    const double* src = e + (si * (JM + 1) + sj) * TS;
                    if (t <= smax + 1) {
                        if (t >= 1) v += oo2p * src[t - 1];
                        if (t <= smax) v += X * src[t];
                        if (t + 1 <= smax) v += (double)(t + 1) * src[t + 1];
                    }
                }
                e[(i * (JM + 1) + j) * TS + t] = v;
            }
#pragma unroll
    for (int k = 0; k < SZ; ++k) E[k] = e[k];
}

// R_tuv (t+u+v <= LR) in registers from the scaled Boys values
// F[n] = pref (-2 alpha)^n F_n(T): R^n_{t..} = (t-1) R^{n+1}_{t-2..} + X R^{n+1}_{t-1..}
// (then u, then v), IntegralsEvaluation.hs:430-449.
template <int LR>
__device__ __forceinline__ void herm_R_regs(const double* F, double X, double Y, double Z, double* R)
{
    double cur[nherm(LR)];
    cur[0] = F[LR];
#pragma unroll
    for (int n = LR - 1; n >= 0; --n) {
        double nxt[nherm(LR)];
        const int top = LR - n;
#pragma unroll
        for (int t = 0; t <= LR; ++t)
#pragma unroll
            for (int u = 0; u <= LR - t; ++u)
#pragma unroll
                for (int v = 0; v <= LR - t - u; ++v) {
                    if (t + u + v > top) continue;
                    double val;
                    if (t > 0) val = (t >= 2 ? (double)(t - 1) * cur[hidx(t - 2, u, v)] : 0.0) + X * cur[hidx(t - 1, u, v)];
                    else if (u > 0) val = (u >= 2 ? (double)(u - 1) * cur[hidx(t, u - 2, v)] : 0.0) + Y * cur[hidx(t, u - 1, v)];
                    else if (v > 0) val = (v >= 2 ? (double)(v - 1) * cur[hidx(t, u, v - 2)] : 0.0) + Z * cur[hidx(t, u, v - 1)];
                    else val = F[n];
                    nxt[hidx(t, u, v)] = val;
                }
#pragma unroll
        for (int e = 0; e < nherm(LR); ++e)
            if (e < nherm(top)) cur[e] = nxt[e];
    }
#pragma unroll
    for (int e = 0; e < nherm(LR); ++e) R[e] = cur[e];
}

template <int LA, int LB, int LC, int LD, int OP>
struct CoopShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NBRA = NA * NB, NKET = NC * ND;
    static constexpr int G = (OP == HFK_OP_GRAD) ? 1 : 0;
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int LR = LAB + LCD + G;          // highest Boys order / Hermite rank
    static constexpr int NH = nherm(LR);
    static constexpr int NX = nherm(LCD + G);         // X^{ab}_{tau nu phi}
    static constexpr int NZ = nherm(LCD);
    static constexpr int S = (OP == HFK_OP_JK || OP == HFK_OP_GRAD) ? 64 / NBRA : 1;
    static constexpr int EB_I = LA + 1 + G, EB_J = LB + 1 + G, EB_T = LAB + 1 + 2 * G;
    static constexpr int EB_SZ = EB_I * EB_J * EB_T;
    static constexpr int EC_I = LC + 1 + G, EC_J = LD + 1, EC_T = LCD + 1 + G;
    static constexpr int EC_SZ = EC_I * EC_J * EC_T;
    static constexpr int NBP_RAW = 768 / (3 * EB_SZ);
    static constexpr int NBP = NBP_RAW < 1 ? 1 : (NBP_RAW > 16 ? 16 : NBP_RAW);   // bra primitive pairs staged at once
    static constexpr bool REG = LR <= 6;   // R_tuv in lane registers (else built level by level in LDS)
};

// LDS of one workgroup
template <class SH>
struct CoopLds {
    double eb[SH::NBP][3][SH::EB_SZ];   // bra Hermite tables per staged primitive pair
    double bs[SH::NBP][8];              // bra primitive pair: P(3), p, K, alpha, qoff
    double ec[SH::S][3][SH::EC_SZ];     // ket Hermite tables of the current ket primitive, per slot
    double ks[SH::S][8];                // ket primitive pair: Q(3), q, K, gamma, qoff
    double R[SH::S][2][SH::NH];         // Hermite integrals, two levels
    double kq[SH::S][SH::NKET];         // ket per-component screen values
    double jcd[SH::S][SH::NKET];        // J'_cd partial sums
    // gradient: the ket's X_CD-derivative coefficients, dE^{cd}_t = 2 gamma E^{c+1,d}_t
    // - c E^{c-1,d}_t - (gamma/q) E^{cd}_{t-1}, indexed like ec
    double dc[SH::G ? SH::S : 1][3][SH::G ? SH::EC_SZ : 1];
};

// Stage the Hermite tables and scalars of bra primitive pairs [bc0, bc0+nbc).
template <class SH>
__device__ __forceinline__ void coop_stage_bra(CoopLds<SH>& sm, const double* bra0, int bc0, int nbc)
{
    const int lane = threadIdx.x;
    if (lane < 3 * nbc) {
        const int ip = lane / 3, ax = lane - 3 * ip;
        const double* r = bra0 + (size_t)(bc0 + ip) * PP_STRIDE;
        herm_table<SH::EB_I - 1, SH::EB_J - 1>(r[PP_AX + ax], r[PP_BX + ax], r[PP_OO2P], &sm.eb[ip][ax][0]);
    }
    if (lane < nbc) {
        const double* r = bra0 + (size_t)(bc0 + lane) * PP_STRIDE;
        sm.bs[lane][0] = r[PP_PX];
        sm.bs[lane][1] = r[PP_PY];
        sm.bs[lane][2] = r[PP_PZ];
        sm.bs[lane][3] = r[PP_P];
        sm.bs[lane][4] = r[PP_K];
        sm.bs[lane][5] = r[PP_EA];
        sm.bs[lane][6] = r[PP_QOFF];
    }
}

// One group of quartets (bra pair bp, this slot's ket kp or -1) over the
// primitive ranges [b0,b1) x [k0,k1).  OP != GRAD: V[cd] of the lane's bra
// component.  GRAD: ga, gb, gc (d/dA, d/dB, d/dC) of sum_cd Gw[cd] (ab|cd).
// staged: the bra tables of [b0,b1) (at most NBP pairs) are already in LDS.
template <int LA, int LB, int LC, int LD, int BOYS, int MODE, int OP>
__device__ __forceinline__ void coop_group(const QuartetArgs& a, CoopLds<CoopShape<LA, LB, LC, LD, OP>>& sm, int bp,
                                           int kp, int b0, int b1, int k0, int k1, bool staged, double* V,
                                           const double* Gw, double* ga, double* gb, double* gc)
{
    using SH = CoopShape<LA, LB, LC, LD, OP>;
    constexpr int NB = SH::NB, ND = SH::ND, NBRA = SH::NBRA, NKET = SH::NKET;
    constexpr int LR = SH::LR, G = SH::G, LAB = SH::LAB, LCD = SH::LCD;
    const int lane = threadIdx.x;
    const int slot = lane / NBRA, li = lane - slot * NBRA;
    const bool active = slot < SH::S;
    const bool has = active && kp >= 0;
    const int ab = active ? li : 0;
    const int ca = ab / NB, cb = ab - (ab / NB) * NB;
    int ea[3], eb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ea[k] = 0;
        eb[k] = 0;
    }
    for (int k = 0; k < 3; ++k) {
        int c = 0;
        for (int i = LA; i >= 0; --i)
            for (int j = LA - i; j >= 0; --j, ++c)
                if (c == ca) ea[k] = (k == 0 ? i : (k == 1 ? j : LA - i - j));
        c = 0;
        for (int i = LB; i >= 0; --i)
            for (int j = LB - i; j >= 0; --j, ++c)
                if (c == cb) eb[k] = (k == 0 ? i : (k == 1 ? j : LB - i - j));
    }
    if constexpr (OP != HFK_OP_GRAD) {
#pragma unroll
        for (int z = 0; z < NKET; ++z) V[z] = 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ga[k] = 0.0;
            gb[k] = 0.0;
            gc[k] = 0.0;
        }
    }
    const double* bra0 = a.pp + (size_t)a.pair_pp0[bp] * PP_STRIDE;
    const double* ket0 = has ? a.pp + (size_t)a.pair_pp0[kp] * PP_STRIDE : nullptr;
    for (int bc0 = b0; bc0 < b1; bc0 += SH::NBP) {
        const int nbc = min(SH::NBP, b1 - bc0);
        if (!staged) {
            __syncthreads();
            coop_stage_bra(sm, bra0, bc0, nbc);
        }
        for (int ik = k0; ik < k1; ++ik) {
            __syncthreads();
            if (has) {
                const double* r = ket0 + (size_t)ik * PP_STRIDE;
                if (li < 3)
                    herm_table<SH::EC_I - 1, SH::EC_J - 1>(r[PP_AX + li], r[PP_BX + li], r[PP_OO2P], &sm.ec[slot][li][0]);
                if constexpr (OP == HFK_OP_GRAD) {
                    if (li < 3) {
                        const double* E = &sm.ec[slot][li][0];
                        double* dE = &sm.dc[slot][li][0];
                        const double gam2 = 2.0 * r[PP_EA], goq = r[PP_EA] / r[PP_P];
                        for (int i = 0; i <= LC; ++i)
                            for (int j = 0; j <= LD; ++j)
                                for (int t = 0; t < SH::EC_T; ++t) {
                                    const int o = (i * SH::EC_J + j) * SH::EC_T;
                                    double v = 0.0;
                                    if (t <= i + j) {
                                        v = gam2 * E[o + SH::EC_J * SH::EC_T + t];
                                        if (i > 0) v -= (double)i * E[o - SH::EC_J * SH::EC_T + t];
                                        if (t >= 1) v -= goq * E[o + t - 1];
                                    }
                                    dE[o + t] = v;
                                }
                    }
                }
                if (li == 3) {
                    sm.ks[slot][0] = r[PP_PX];
                    sm.ks[slot][1] = r[PP_PY];
                    sm.ks[slot][2] = r[PP_PZ];
                    sm.ks[slot][3] = r[PP_P];
                    sm.ks[slot][4] = r[PP_K];
                    sm.ks[slot][5] = r[PP_EA];
                }
                if constexpr (MODE == 1) {
                    const int qoff = (int)r[PP_QOFF];
                    for (int z = li; z < NKET; z += NBRA) sm.kq[slot][z] = a.qcomp[qoff + z];
                }
            }
            __syncthreads();
            for (int ib = 0; ib < nbc; ++ib) {
                if constexpr (SH::REG) {
                    if (has) {
                        const double* Eb = &sm.eb[ib][0][0];
                        const double* Ec = &sm.ec[slot][0][0];
                        auto EB = [&](int axis, int i, int j, int t) -> double {
                            return Eb[axis * SH::EB_SZ + (i * SH::EB_J + j) * SH::EB_T + t];
                        };
                        auto EC = [&](int axis, int i, int j, int t) -> double {
                            return Ec[axis * SH::EC_SZ + (i * SH::EC_J + j) * SH::EC_T + t];
                        };
                        double R[SH::NH];
                        {
                            double F[LR + 1];
                            const double p = sm.bs[ib][3], q = sm.ks[slot][3];
                            const double pq = p + q, alpha = p * q / pq;
                            const double X = sm.bs[ib][0] - sm.ks[slot][0];
                            const double Y = sm.bs[ib][1] - sm.ks[slot][1];
                            const double Z = sm.bs[ib][2] - sm.ks[slot][2];
                            boys_eval<double, LR, BOYS>(alpha * ((X * X + Y * Y) + Z * Z), a.bt, F);
                            double w = HFK_TWO_PI_2_5 * sm.bs[ib][4] * sm.ks[slot][4] / sqrt(pq);
#pragma unroll
                            for (int n = 0; n <= LR; ++n) {
                                F[n] *= w;
                                w *= -2.0 * alpha;
                            }
                            herm_R_regs<LR>(F, X, Y, Z, R);
                        }
                        // the lane's bra component, dense over t+u+v <= LAB (+1): zero beyond a+b
                        constexpr int LE = LAB + 1 + G;
                        double e1[3][LE];
#pragma unroll
                        for (int x = 0; x < 3; ++x)
#pragma unroll
                            for (int t = 0; t < LE; ++t) e1[x][t] = (t < SH::EB_T) ? EB(x, ea[x], eb[x], t) : 0.0;
                        // X_{tau nu phi} = sum_tuv E^{ab}_tuv R_{t+tau,u+nu,v+phi}, tau+nu+phi <= LCD
                        double Xv[SH::NX];
#pragma unroll
                        for (int e = 0; e < SH::NX; ++e) Xv[e] = 0.0;
#pragma unroll
                        for (int t = 0; t <= LAB; ++t)
#pragma unroll
                            for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
                                for (int v = 0; v <= LAB - t - u; ++v) {
                                    const double c = (e1[0][t] * e1[1][u]) * e1[2][v];
#pragma unroll
                                    for (int tau = 0; tau <= LCD; ++tau)
#pragma unroll
                                        for (int nu = 0; nu <= LCD - tau; ++nu)
#pragma unroll
                                            for (int ph = 0; ph <= LCD - tau - nu; ++ph)
                                                Xv[hidx(tau, nu, ph)] += c * R[hidx(t + tau, u + nu, v + ph)];
                                }
                        double bq = 0.0;
                        if constexpr (MODE == 1) bq = a.qcomp[(int)sm.bs[ib][6] + ab];
                        if constexpr (OP != HFK_OP_GRAD) {
#pragma unroll
                            for (int cd = 0; cd < NKET; ++cd) {
                                const int c = cd / ND, d = cd - (cd / ND) * ND;
                                const int cx = cart_e(LC, c, 0), cy = cart_e(LC, c, 1), cz = cart_e(LC, c, 2);
                                const int dx = cart_e(LD, d, 0), dy = cart_e(LD, d, 1), dz = cart_e(LD, d, 2);
                                double s = 0.0;
#pragma unroll
                                for (int tau = 0; tau <= cx + dx; ++tau)
#pragma unroll
                                    for (int nu = 0; nu <= cy + dy; ++nu) {
                                        const double exy = EC(0, cx, dx, tau) * EC(1, cy, dy, nu);
#pragma unroll
                                        for (int ph = 0; ph <= cz + dz; ++ph) {
                                            const double term = exy * EC(2, cz, dz, ph) * Xv[hidx(tau, nu, ph)];
                                            s += ((tau + nu + ph) & 1) ? -term : term;
                                        }
                                    }
                                if constexpr (MODE == 1) V[cd] += !(bq * sm.kq[slot][cd] < a.qtau) ? s : 0.0;
                                else V[cd] += s;
                            }
                        } else {
                            // dA from the shifted bra coefficients; dA + dB = dP, the contraction
                            // with R shifted by one along the axis; dC = -(gamma/q) dP plus the
                            // ket's X_CD-derivative coefficients (dc); dD by translational
                            // invariance.  Zv: the ket coefficients contracted with the density,
                            // Zd[k]: the same with the derivative along k.
                            double Zv[SH::NZ], Zd[3][SH::NZ];
#pragma unroll
                            for (int e = 0; e < SH::NZ; ++e) {
                                Zv[e] = 0.0;
                                Zd[0][e] = 0.0;
                                Zd[1][e] = 0.0;
                                Zd[2][e] = 0.0;
                            }
                            const double* Dc = &sm.dc[slot][0][0];
                            auto DC = [&](int axis, int i, int j, int t) -> double {
                                return Dc[axis * SH::EC_SZ + (i * SH::EC_J + j) * SH::EC_T + t];
                            };
#pragma unroll
                            for (int cd = 0; cd < NKET; ++cd) {
                                const int c = cd / ND, d = cd - (cd / ND) * ND;
                                const int cc[3] = {cart_e(LC, c, 0), cart_e(LC, c, 1), cart_e(LC, c, 2)};
                                const int dd[3] = {cart_e(LD, d, 0), cart_e(LD, d, 1), cart_e(LD, d, 2)};
                                double w = Gw[cd];
                                if constexpr (MODE == 1) w = !(bq * sm.kq[slot][cd] < a.qtau) ? w : 0.0;
                                double e[3][LCD + 1], de[3][LCD + 1];
#pragma unroll
                                for (int x = 0; x < 3; ++x)
#pragma unroll
                                    for (int t = 0; t <= LCD; ++t) {
                                        e[x][t] = t <= cc[x] + dd[x] ? EC(x, cc[x], dd[x], t) : 0.0;
                                        de[x][t] = t <= cc[x] + dd[x] ? DC(x, cc[x], dd[x], t) : 0.0;
                                    }
#pragma unroll
                                for (int tau = 0; tau <= cc[0] + dd[0]; ++tau)
#pragma unroll
                                    for (int nu = 0; nu <= cc[1] + dd[1]; ++nu)
#pragma unroll
                                        for (int ph = 0; ph <= cc[2] + dd[2]; ++ph) {
                                            const double ws = ((tau + nu + ph) & 1) ? -w : w;
                                            const int h = hidx(tau, nu, ph);
                                            const double eyz = e[1][nu] * e[2][ph], exz = e[0][tau] * e[2][ph];
                                            const double exy = e[0][tau] * e[1][nu];
                                            Zv[h] += ws * (e[0][tau] * eyz);
                                            Zd[0][h] += ws * (de[0][tau] * eyz);
                                            Zd[1][h] += ws * (de[1][nu] * exz);
                                            Zd[2][h] += ws * (de[2][ph] * exy);
                                        }
                            }
                            double scd[3] = {0.0, 0.0, 0.0};
#pragma unroll
                            for (int e = 0; e < SH::NZ; ++e) {
                                scd[0] += Zd[0][e] * Xv[e];
                                scd[1] += Zd[1][e] * Xv[e];
                                scd[2] += Zd[2][e] * Xv[e];
                            }
                            // Y_tuv = sum_{tau nu phi} Z R_{t+tau,...}, t+u+v <= LAB+1
                            double Yv[nherm(LAB + 1)];
#pragma unroll
                            for (int t = 0; t <= LAB + 1; ++t)
#pragma unroll
                                for (int u = 0; u <= LAB + 1 - t; ++u)
#pragma unroll
                                    for (int v = 0; v <= LAB + 1 - t - u; ++v) {
                                        double y = 0.0;
#pragma unroll
                                        for (int tau = 0; tau <= LCD; ++tau)
#pragma unroll
                                            for (int nu = 0; nu <= LCD - tau; ++nu)
#pragma unroll
                                                for (int ph = 0; ph <= LCD - tau - nu; ++ph)
                                                    y += Zv[hidx(tau, nu, ph)] * R[hidx(t + tau, u + nu, v + ph)];
                                        Yv[hidx(t, u, v)] = y;
                                    }
                            // dA coefficients of the lane's component (dense; zero beyond a+b+1)
                            const double alp2 = 2.0 * sm.bs[ib][5];
                            double dA1[3][LE];
#pragma unroll
                            for (int x = 0; x < 3; ++x) {
                                const int i = ea[x], j = eb[x];
#pragma unroll
                                for (int t = 0; t < LE; ++t) {
                                    double va = alp2 * EB(x, i + 1, j, t);
                                    if (i > 0) va -= (double)i * EB(x, i - 1, j, t);
                                    dA1[x][t] = va;
                                }
                            }
                            // factorised sums over v, then u, then t:
                            //   gA_x = sum dA_x(t) E_y(u) E_z(v) Y_tuv, and likewise y, z;
                            //   dP_x = sum E_x(t) E_y(u) E_z(v) Y_{t+1,u,v}, and likewise y, z
                            double gq[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
#pragma unroll
                            for (int t = 0; t <= LAB + 1; ++t) {
                                double At = 0.0, Bt = 0.0, Ct = 0.0, Dt = 0.0, Et = 0.0;
#pragma unroll
                                for (int u = 0; u <= LAB + 1 - t; ++u) {
                                    double pz = 0.0, qz = 0.0, sz = 0.0;
#pragma unroll
                                    for (int v = 0; v <= LAB + 1 - t - u; ++v) {
                                        const double y = Yv[hidx(t, u, v)];
                                        pz += e1[2][v] * y;
                                        qz += dA1[2][v] * y;
                                        if (v >= 1) sz += e1[2][v - 1] * y;
                                    }
                                    At += e1[1][u] * pz;
                                    Bt += dA1[1][u] * pz;
                                    Ct += e1[1][u] * qz;
                                    if (u >= 1) Dt += e1[1][u - 1] * pz;
                                    Et += e1[1][u] * sz;
                                }
                                gq[0] += dA1[0][t] * At;
                                gq[1] += e1[0][t] * Bt;
                                gq[2] += e1[0][t] * Ct;
                                if (t >= 1) sq[0] += e1[0][t - 1] * At;
                                sq[1] += e1[0][t] * Dt;
                                sq[2] += e1[0][t] * Et;
                            }
                            const double goq = sm.ks[slot][5] / sm.ks[slot][3];
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                ga[k] += gq[k];
                                gb[k] += sq[k] - gq[k];
                                gc[k] += scd[k] - goq * sq[k];
                            }
                        }
                    }
                } else {
                {
                // Boys values and R^n_000, then R_tuv level by level
                double F[LR + 1];
                {
                    const double p = sm.bs[ib][3];
                    const double q = has ? sm.ks[slot][3] : 1.0;
                    const double pq = p + q, alpha = p * q / pq;
                    const double X = has ? sm.bs[ib][0] - sm.ks[slot][0] : 0.0;
                    const double Y = has ? sm.bs[ib][1] - sm.ks[slot][1] : 0.0;
                    const double Z = has ? sm.bs[ib][2] - sm.ks[slot][2] : 0.0;
                    const double T = alpha * ((X * X + Y * Y) + Z * Z);
                    boys_eval<double, LR, BOYS>(T, a.bt, F);
                    const double pref = has ? HFK_TWO_PI_2_5 * sm.bs[ib][4] * sm.ks[slot][4] / sqrt(pq) : 0.0;
                    double w = pref;
#pragma unroll
                    for (int n = 0; n <= LR; ++n) {
                        F[n] *= w;
                        w *= -2.0 * alpha;
                    }
                    double* Rb = &sm.R[slot][0][0];
#pragma unroll
                    for (int n = LR; n >= 0; --n) {
                        double* cur = Rb + (n & 1) * SH::NH;
                        const double* prv = Rb + ((n + 1) & 1) * SH::NH;
                        if (has) {
                            for (int e = li; e < nherm(LR - n); e += NBRA) {
                                int t, u, v;
                                hdecode(e, t, u, v);
                                double val;
                                if (t > 0) val = (t >= 2 ? (double)(t - 1) * prv[hidx(t - 2, u, v)] : 0.0) + X * prv[hidx(t - 1, u, v)];
                                else if (u > 0) val = (u >= 2 ? (double)(u - 1) * prv[hidx(t, u - 2, v)] : 0.0) + Y * prv[hidx(t, u - 1, v)];
                                else if (v > 0) val = (v >= 2 ? (double)(v - 1) * prv[hidx(t, u, v - 2)] : 0.0) + Z * prv[hidx(t, u, v - 1)];
                                else val = F[n];
                                cur[e] = val;
                            }
                        }
                        __syncthreads();
                    }
                }
                if (has) {
                const double* R = &sm.R[slot][0][0];
                const double* Eb = &sm.eb[ib][0][0];
                const double* Ec = &sm.ec[slot][0][0];
                auto EB = [&](int axis, int i, int j, int t) -> double {
                    return Eb[axis * SH::EB_SZ + (i * SH::EB_J + j) * SH::EB_T + t];
                };
                auto EC = [&](int axis, int i, int j, int t) -> double {
                    return Ec[axis * SH::EC_SZ + (i * SH::EC_J + j) * SH::EC_T + t];
                };
                // X^{ab}_{tau nu phi} = sum_tuv E^{ab}_tuv R_{t+tau,u+nu,v+phi}
                double Xv[SH::NX];
                {
                    double ex[SH::LAB + 1], ey[SH::LAB + 1], ez[SH::LAB + 1];
#pragma unroll
                    for (int t = 0; t <= SH::LAB; ++t) {
                        ex[t] = t <= ea[0] + eb[0] ? EB(0, ea[0], eb[0], t) : 0.0;
                        ey[t] = t <= ea[1] + eb[1] ? EB(1, ea[1], eb[1], t) : 0.0;
                        ez[t] = t <= ea[2] + eb[2] ? EB(2, ea[2], eb[2], t) : 0.0;
                    }
#pragma unroll
                    for (int e = 0; e < SH::NX; ++e) Xv[e] = 0.0;
                    for (int t = 0; t <= ea[0] + eb[0]; ++t)
                        for (int u = 0; u <= ea[1] + eb[1]; ++u)
                            for (int v = 0; v <= ea[2] + eb[2]; ++v) {
                                const double c = (ex[t] * ey[u]) * ez[v];
#pragma unroll
                                for (int n = 0; n <= SH::LCD + G; ++n)
#pragma unroll
                                    for (int m = 0; m <= n; ++m)
#pragma unroll
                                        for (int ph = 0; ph <= m; ++ph) {
                                            const int tau = n - m, nu = m - ph;
                                            Xv[hidx(tau, nu, ph)] += c * R[hidx(t + tau, u + nu, v + ph)];
                                        }
                            }
                }
                double bq = 0.0;
                if constexpr (MODE == 1) bq = a.qcomp[(int)sm.bs[ib][6] + ab];
                if constexpr (OP != HFK_OP_GRAD) {
#pragma unroll
                    for (int cd = 0; cd < NKET; ++cd) {
                        const int c = cd / ND, d = cd - (cd / ND) * ND;
                        const int cx = cart_e(LC, c, 0), cy = cart_e(LC, c, 1), cz = cart_e(LC, c, 2);
                        const int dx = cart_e(LD, d, 0), dy = cart_e(LD, d, 1), dz = cart_e(LD, d, 2);
                        double s = 0.0;
#pragma unroll
                        for (int tau = 0; tau <= cx + dx; ++tau)
#pragma unroll
                            for (int nu = 0; nu <= cy + dy; ++nu) {
                                const double exy = EC(0, cx, dx, tau) * EC(1, cy, dy, nu);
#pragma unroll
                                for (int ph = 0; ph <= cz + dz; ++ph) {
                                    const double term = exy * EC(2, cz, dz, ph) * Xv[hidx(tau, nu, ph)];
                                    s += ((tau + nu + ph) & 1) ? -term : term;
                                }
                            }
                        if constexpr (MODE == 1) V[cd] += !(bq * sm.kq[slot][cd] < a.qtau) ? s : 0.0;
                        else V[cd] += s;
                    }
                } else {
                    // density-weighted ket coefficients Z_{tau nu phi} = sum_cd Gw (-1)^{..} E^{cd}
                    double Zv[SH::NZ];
#pragma unroll
                    for (int e = 0; e < SH::NZ; ++e) Zv[e] = 0.0;
                    double gck[3] = {0.0, 0.0, 0.0};
                    const double gam2 = 2.0 * sm.ks[slot][5];
#pragma unroll
                    for (int cd = 0; cd < NKET; ++cd) {
                        const int c = cd / ND, d = cd - (cd / ND) * ND;
                        const int cc[3] = {cart_e(LC, c, 0), cart_e(LC, c, 1), cart_e(LC, c, 2)};
                        const int dd[3] = {cart_e(LD, d, 0), cart_e(LD, d, 1), cart_e(LD, d, 2)};
                        double w = Gw[cd];
                        if constexpr (MODE == 1) w = !(bq * sm.kq[slot][cd] < a.qtau) ? w : 0.0;
#pragma unroll
                        for (int tau = 0; tau <= cc[0] + dd[0]; ++tau)
#pragma unroll
                            for (int nu = 0; nu <= cc[1] + dd[1]; ++nu)
#pragma unroll
                                for (int ph = 0; ph <= cc[2] + dd[2]; ++ph) {
                                    const double e = EC(0, cc[0], dd[0], tau) * EC(1, cc[1], dd[1], nu) * EC(2, cc[2], dd[2], ph);
                                    Zv[hidx(tau, nu, ph)] += ((tau + nu + ph) & 1) ? -w * e : w * e;
                                }
                        // d/dC: 2 gamma E^{c+1,d} - c E^{c-1,d} along axis k
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            double s = 0.0;
#pragma unroll
                            for (int tau = 0; tau <= cc[0] + dd[0] + (k == 0); ++tau)
#pragma unroll
                                for (int nu = 0; nu <= cc[1] + dd[1] + (k == 1); ++nu)
#pragma unroll
                                    for (int ph = 0; ph <= cc[2] + dd[2] + (k == 2); ++ph) {
                                        const int idx[3] = {tau, nu, ph};
                                        double f = 1.0;
#pragma unroll
                                        for (int x = 0; x < 3; ++x) {
                                            if (x == k) {
                                                double dv = gam2 * EC(x, cc[x] + 1, dd[x], idx[x]);
                                                if (cc[x] > 0 && idx[x] <= cc[x] - 1 + dd[x])
                                                    dv -= (double)cc[x] * EC(x, cc[x] - 1, dd[x], idx[x]);
                                                f *= dv;
                                            } else {
                                                f *= EC(x, cc[x], dd[x], idx[x]);
                                            }
                                        }
                                        const double term = f * Xv[hidx(tau, nu, ph)];
                                        s += ((tau + nu + ph) & 1) ? -term : term;
                                    }
                            gck[k] += w * s;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) gc[k] += gck[k];
                    // d/dA, d/dB: Y_tuv = sum Z R_{t+tau,...} over the derivative boxes
                    const double alp2 = 2.0 * sm.bs[ib][5];
                    const double bet2 = 2.0 * (sm.bs[ib][3] - sm.bs[ib][5]);
                    for (int k = 0; k < 3; ++k) {
                        double dA[SH::LAB + 2], dB[SH::LAB + 2];
                        const int i = ea[k], j = eb[k];
                        for (int t = 0; t <= i + j + 1; ++t) {
                            double va = alp2 * EB(k, i + 1, j, t);
                            if (i > 0 && t <= i + j - 1) va -= (double)i * EB(k, i - 1, j, t);
                            double vb = bet2 * EB(k, i, j + 1, t);
                            if (j > 0 && t <= i + j - 1) vb -= (double)j * EB(k, i, j - 1, t);
                            dA[t] = va;
                            dB[t] = vb;
                        }
                        const int k1 = k == 0 ? 1 : 0, k2 = k == 2 ? 1 : 2;
                        double sa = 0.0, sb = 0.0;
                        for (int t = 0; t <= i + j + 1; ++t)
                            for (int u = 0; u <= ea[k1] + eb[k1]; ++u)
                                for (int v = 0; v <= ea[k2] + eb[k2]; ++v) {
                                    int h[3];
                                    h[k] = t;
                                    h[k1] = u;
                                    h[k2] = v;
                                    double y = 0.0;
#pragma unroll
                                    for (int n = 0; n <= SH::LCD; ++n)
#pragma unroll
                                        for (int m = 0; m <= n; ++m)
#pragma unroll
                                            for (int ph = 0; ph <= m; ++ph) {
                                                const int tau = n - m, nu = m - ph;
                                                y += Zv[hidx(tau, nu, ph)] * R[hidx(h[0] + tau, h[1] + nu, h[2] + ph)];
                                            }
                                    const double eo = EB(k1, ea[k1], eb[k1], u) * EB(k2, ea[k2], eb[k2], v);
                                    sa += dA[t] * eo * y;
                                    sb += dB[t] * eo * y;
                                }
                        ga[k] += sa;
                        gb[k] += sb;
                    }
                }
                }   // has
                }
                }
                if constexpr (!SH::REG) __syncthreads();   // R and the staged tables are rewritten next
            }
        }
    }
}

template <int LA, int LB, int LC, int LD, int BOYS, int MODE, int OP>
__global__ void __launch_bounds__(64) k_coop(QuartetArgs a)
{
    using SH = CoopShape<LA, LB, LC, LD, OP>;
    constexpr int NA = SH::NA, NB = SH::NB, NC = SH::NC, ND = SH::ND, NBRA = SH::NBRA, NKET = SH::NKET;
    __shared__ CoopLds<SH> sm;
    extern __shared__ double dyn[];   // JK: K rows of the bra shells; GRAD: atom accumulator
    const int lane = threadIdx.x;
    const int slot = lane / NBRA, li = lane - slot * NBRA;
    const bool active = slot < SH::S;
    const int ab = active ? li : 0;
    const int ca = ab / NB, cb = ab - (ab / NB) * NB;
    const int ld = a.n + 1;

    if constexpr (OP == HFK_OP_FLAT || OP == HFK_OP_DIAG) {
        double V[NKET];
        for (long long q = blockIdx.x; q < a.nquart; q += gridDim.x) {
            int bp, kp;
            if (OP == HFK_OP_DIAG) {
                bp = kp = a.bra_list[q];
            } else {
                long long i, j;
                decode_quartet(a, q, i, j);
                bp = a.bra_list[i];
                kp = a.ket_list[j];
                if (!quartet_selected(a, bp, kp)) continue;   // uniform over the workgroup
            }
            coop_group<LA, LB, LC, LD, BOYS, MODE, OP>(a, sm, bp, active ? kp : -1, 0, a.pair_npp[bp], 0,
                                                       a.pair_npp[kp], false, V, nullptr, nullptr, nullptr, nullptr);
            const int A = a.pair_A[bp], B = a.pair_B[bp], C = a.pair_A[kp], Dd = a.pair_B[kp];
            if constexpr (OP == HFK_OP_FLAT) {
                if (active) {
                    const int i = a.sh_bf[A * HFK_MAXCOMP + ca], j = a.sh_bf[B * HFK_MAXCOMP + cb];
                    const double sij = a.sh_scale[A * HFK_MAXCOMP + ca] * a.sh_scale[B * HFK_MAXCOMP + cb];
#pragma unroll
                    for (int cd = 0; cd < NKET; ++cd) {
                        const int c = cd / ND, d = cd - (cd / ND) * ND;
                        const int k = a.sh_bf[C * HFK_MAXCOMP + c], l = a.sh_bf[Dd * HFK_MAXCOMP + d];
                        if (i < 0 || j < 0 || k < 0 || l < 0) continue;
                        const double s = sij * a.sh_scale[C * HFK_MAXCOMP + c] * a.sh_scale[Dd * HFK_MAXCOMP + d];
                        a.flat[flat_pos(a, i, j, k, l)] = V[cd] * s;
                    }
                }
            } else {
                double m = 0.0;
                if (active) {
                    const double s = a.sh_scale[A * HFK_MAXCOMP + ca] * a.sh_scale[B * HFK_MAXCOMP + cb];
                    double vd = 0.0;
#pragma unroll
                    for (int cd = 0; cd < NKET; ++cd)
                        if (cd == ab) vd = V[cd];
                    m = fabs(vd) * s * s;
                }
#pragma unroll
                for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, WAVE));
                if (lane == 0) a.diag[bp] = m;
            }
        }
        return;
    } else if constexpr (OP == HFK_OP_JK) {
        // J/K: a task is one run of kets with explicit primitive ranges
        // (uniform over the workgroup), walked S kets at a time.  One wave,
        // so the LDS accumulators (K' rows of the bra shells, the slots'
        // J'_cd sums) take the lanes' adds in a fixed hardware order; the
        // results go to the task's slots with plain stores.
        const long long tix = a.task0 + blockIdx.x;
        const int4 task = a.tasks[blockIdx.x];
        const int len = task_ket_count(a, task);
        const int k0 = (task.w >> 8) & 0xff, k1 = task.w & 0xff;
        // uniform over the workgroup, before any barrier; this kernel needs an
        // explicit ket primitive range (the planner never gives it k1 = all)
        if (k1 == 0xff) return;
        const int bp = task.x;
        const int b0 = (task.w >> 24) & 0xff, b1 = (task.w >> 16) & 0xff;
        const int A = a.pair_A[bp], B = a.pair_B[bp];
        const int ia = a.sh_bfp[A * HFK_MAXCOMP + ca], ib = a.sh_bfp[B * HFK_MAXCOMP + cb];
        const double sab = active ? a.sh_scale[A * HFK_MAXCOMP + ca] * a.sh_scale[B * HFK_MAXCOMP + cb] : 0.0;
        const double* __restrict__ D = a.D;
        const double dab = D[(size_t)ia * ld + ib];
        const int ldk = a.ncols + 1;
        double* krow = dyn;                    // [NA+NB][ldk]
        double* sred = dyn + (NA + NB) * ldk;  // [64] per-lane J'_ab
        for (int t = lane; t < (NA + NB) * ldk; t += 64) krow[t] = 0.0;
        for (int t = lane; t < SH::S * NKET; t += 64) (&sm.jcd[0][0])[t] = 0.0;
        // bra tables once per task when its primitive range fits the staging area
        const bool staged = (b1 - b0) <= SH::NBP;
        if (staged) coop_stage_bra(sm, a.pp + (size_t)a.pair_pp0[bp] * PP_STRIDE, b0, b1 - b0);
        __syncthreads();
        const KetRun run = a.runs[task.y];
        const int* __restrict__ colmap = a.colmap;
        double jab = 0.0;
        for (int g0 = 0; g0 < len; g0 += SH::S) {
            const int it = g0 + slot;
            int kp = -1;
            if (active && it < len) {
                kp = a.ket_idx[run.x + it];
                if (!quartet_selected(a, bp, kp)) kp = -1;
            }
            int C = 0, Dd = 0;
            int ic[NC], id[ND], cc[NC], cdc[ND];
            double sc[NC], sd[ND];
            double f = 0.0;
            if (kp >= 0) {
                C = a.pair_A[kp];
                Dd = a.pair_B[kp];
                f = 1.0;
                if (A == B) f *= 0.5;
                if (C == Dd) f *= 0.5;
                if (bp == kp) f *= 0.5;
            }
#pragma unroll
            for (int x = 0; x < NC; ++x) {
                ic[x] = kp >= 0 ? a.sh_bfp[C * HFK_MAXCOMP + x] : a.n;
                sc[x] = kp >= 0 ? a.sh_scale[C * HFK_MAXCOMP + x] : 0.0;
                cc[x] = colmap[ic[x]];
            }
#pragma unroll
            for (int x = 0; x < ND; ++x) {
                id[x] = kp >= 0 ? a.sh_bfp[Dd * HFK_MAXCOMP + x] : a.n;
                sd[x] = kp >= 0 ? a.sh_scale[Dd * HFK_MAXCOMP + x] : 0.0;
                cdc[x] = colmap[id[x]];
            }
            double V[NKET];
            coop_group<LA, LB, LC, LD, BOYS, MODE, OP>(a, sm, bp, kp, b0, b1, k0, k1, staged, V, nullptr, nullptr,
                                                       nullptr, nullptr);
            if (kp >= 0) {
                const size_t rowa = (size_t)ca * ldk, rowb = (size_t)(NA + cb) * ldk;
                double js = 0.0;
#pragma unroll
                for (int cd = 0; cd < NKET; ++cd) {
                    const int c = cd / ND, d = cd - (cd / ND) * ND;
                    V[cd] *= f * sab * sc[c] * sd[d];
                    js += V[cd] * D[(size_t)ic[c] * ld + id[d]];
                    atomic_add_lds(&sm.jcd[slot][cd], 2.0 * V[cd] * dab);
                }
                jab += 2.0 * js;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    double s1 = 0.0, s3 = 0.0;
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        s1 += V[c * ND + d] * D[(size_t)ib * ld + id[d]];
                        s3 += V[c * ND + d] * D[(size_t)ia * ld + id[d]];
                    }
                    atomic_add_lds(&krow[rowa + cc[c]], s1);
                    atomic_add_lds(&krow[rowb + cc[c]], s3);
                }
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    double s2 = 0.0, s4 = 0.0;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        s2 += V[c * ND + d] * D[(size_t)ib * ld + ic[c]];
                        s4 += V[c * ND + d] * D[(size_t)ia * ld + ic[c]];
                    }
                    atomic_add_lds(&krow[rowa + cdc[d]], s2);
                    atomic_add_lds(&krow[rowb + cdc[d]], s4);
                }
            }
            __syncthreads();
            // every ket of the task rewrites its J'_cd entry (zero when skipped)
            if (active && it < len) {
                for (int z = li; z < NKET; z += NBRA) {
                    a.jcd[(run.jcd + it) * NKET + z] = sm.jcd[slot][z];
                    sm.jcd[slot][z] = 0.0;
                }
            }
            __syncthreads();
        }
        // J'_ab: the slots' partial sums in slot order
        sred[lane] = active ? jab : 0.0;
        __syncthreads();
        if (lane < NBRA) {
            double v = 0.0;
            for (int sl = 0; sl < SH::S; ++sl) v += sred[sl * NBRA + lane];
            a.jab[(size_t)a.jab_slot[tix] * NBRA + lane] = v;
        }
        double* __restrict__ ks = a.kslot + (size_t)tix * (NA + NB) * a.ncols;
        for (int r = 0; r < NA + NB; ++r)
            for (int col = lane; col < a.ncols; col += 64) ks[(size_t)r * a.ncols + col] = krow[r * ldk + col];
    } else {
        const int4 task = a.tasks[blockIdx.x];
        // a long task (planned for the per-lane kernel) is cut into slices of
        // 4 S kets along grid y; every slice adds into the gradient independently
        const int kslice = 4 * SH::S;
        const int kbeg = (int)blockIdx.y * kslice;   // kets of the task, in run order
        const int ktot = task_ket_count(a, task);
        const int k0 = (task.w >> 8) & 0xff, k1 = task.w & 0xff;
        // uniform over the workgroup, before any barrier; this kernel needs an
        // explicit ket primitive range (the planner never gives it k1 = all)
        if (kbeg >= ktot || k1 == 0xff) return;
        const int kend = min(ktot, kbeg + kslice);
        const int bp = task.x;
        const int b0 = (task.w >> 24) & 0xff, b1 = (task.w >> 16) & 0xff;
        const int A = a.pair_A[bp], B = a.pair_B[bp];
        const int ia = a.sh_bfp[A * HFK_MAXCOMP + ca], ib = a.sh_bfp[B * HFK_MAXCOMP + cb];
        const double sab = active ? a.sh_scale[A * HFK_MAXCOMP + ca] * a.sh_scale[B * HFK_MAXCOMP + cb] : 0.0;
        const double* __restrict__ D = a.D;
        const double dab = D[(size_t)ia * ld + ib];
        const bool lds = a.grad_lds != 0;
        const int ndyn = 3 * a.natom;
        if (lds)
            for (int t = lane; t < ndyn; t += 64) dyn[t] = 0.0;
        // bra tables once per task when its primitive range fits the staging area
        const bool staged = (b1 - b0) <= SH::NBP;
        if (staged) coop_stage_bra(sm, a.pp + (size_t)a.pair_pp0[bp] * PP_STRIDE, b0, b1 - b0);
        __syncthreads();
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
        const int len = kend - kbeg;
        for (int g0 = 0; g0 < len; g0 += SH::S) {
            const int it = g0 + slot;
            int kp = -1;
            if (active && it < len) {
                kp = a.ket_idx[task_ket_pos(a, task, kbeg + it)];
                if (!quartet_selected(a, bp, kp)) kp = -1;
            }
            int C = 0, Dd = 0;
            int ic[NC], id[ND];
            double sc[NC], sd[ND];
            double f = 0.0;
            if (kp >= 0) {
                C = a.pair_A[kp];
                Dd = a.pair_B[kp];
                f = 1.0;
                if (A == B) f *= 0.5;
                if (C == Dd) f *= 0.5;
                if (bp == kp) f *= 0.5;
            }
#pragma unroll
            for (int x = 0; x < NC; ++x) {
                ic[x] = kp >= 0 ? a.sh_bfp[C * HFK_MAXCOMP + x] : a.n;
                sc[x] = kp >= 0 ? a.sh_scale[C * HFK_MAXCOMP + x] : 0.0;
            }
#pragma unroll
            for (int x = 0; x < ND; ++x) {
                id[x] = kp >= 0 ? a.sh_bfp[Dd * HFK_MAXCOMP + x] : a.n;
                sd[x] = kp >= 0 ? a.sh_scale[Dd * HFK_MAXCOMP + x] : 0.0;
            }
            // two-particle density of the lane's bra row: 4 f s_abcd Gamma
            double Gw[NKET];
#pragma unroll
            for (int cd = 0; cd < NKET; ++cd) {
                const int c = cd / ND, d = cd - (cd / ND) * ND;
                const double dcd = D[(size_t)ic[c] * ld + id[d]];
                const double dac = D[(size_t)ia * ld + ic[c]], dbd = D[(size_t)ib * ld + id[d]];
                const double dad = D[(size_t)ia * ld + id[d]], dbc = D[(size_t)ib * ld + ic[c]];
                Gw[cd] = (4.0 * f * sab * sc[c] * sd[d]) * (dab * dcd - 0.25 * (dac * dbd + dad * dbc));
            }
            double ga[3], gb[3], gc[3];
            coop_group<LA, LB, LC, LD, BOYS, MODE, OP>(a, sm, bp, kp, b0, b1, k0, k1, staged, nullptr, Gw, ga, gb,
                                                       gc);
            if (kp >= 0) {
                double* acc = lds ? dyn : a.grad;
                const int atC = a.sh_atom[C], atD = a.sh_atom[Dd];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gA[k] += ga[k];
                    gB[k] += gb[k];
                    atomic_add_f64(&acc[3 * atC + k], gc[k]);
                    atomic_add_f64(&acc[3 * atD + k], -((ga[k] + gb[k]) + gc[k]));
                }
            }
        }
        double* acc = lds ? dyn : a.grad;
        const int atA = a.sh_atom[A], atB = a.sh_atom[B];
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
            for (int t = lane; t < 3 * a.natom; t += 64)
                if (dyn[t] != 0.0) atomic_add_f64(&a.grad[t], dyn[t]);
        }
    }
}

// 0, or -3 when the K' rows do not fit in LDS next to the static tables
template <int LA, int LB, int LC, int LD, int BOYS, int MODE, int OP>
int launch_coop(const QuartetArgs& a, hipStream_t s)
{
    if constexpr (OP == HFK_OP_JK) {
        if (a.ntask <= 0) return 0;
        const size_t shm = hfk_coop_jk_lds_bytes(LA, LB, a.ncols);
        if (shm + sizeof(CoopLds<CoopShape<LA, LB, LC, LD, OP>>) > 64 * 1024) return -3;
        long long done = 0;
        while (done < a.ntask) {
            const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
            QuartetArgs b = a;
            b.tasks = a.tasks + done;
            b.task0 = a.task0 + done;
            hipLaunchKernelGGL((k_coop<LA, LB, LC, LD, BOYS, MODE, OP>), dim3((unsigned)cnt), dim3(64), shm, s, b);
            done += cnt;
        }
    } else if constexpr (OP == HFK_OP_GRAD) {
        if (a.ntask <= 0) return 0;
        using SH = CoopShape<LA, LB, LC, LD, OP>;
        QuartetArgs base = a;
        size_t shm = a.grad_lds ? sizeof(double) * 3 * (size_t)a.natom : 0;
        if (shm + sizeof(CoopLds<SH>) > 64 * 1024) {   // stage nothing: global atomics
            shm = 0;
            base.grad_lds = 0;
        }
        long long done = 0;
        while (done < a.ntask) {
            const long long cnt = std::min<long long>(a.ntask - done, 1LL << 30);
            QuartetArgs b = base;
            b.tasks = a.tasks + done;
            const int kslice = 4 * SH::S;
            const unsigned ysl = (unsigned)std::max(1, (a.task_len_max + kslice - 1) / kslice);
            hipLaunchKernelGGL((k_coop<LA, LB, LC, LD, BOYS, MODE, OP>), dim3((unsigned)cnt, ysl), dim3(64), shm, s, b);
            done += cnt;
        }
    } else {
        if (a.nquart <= 0) return 0;
        const long long g = std::min<long long>(a.nquart, 65536);
        hipLaunchKernelGGL((k_coop<LA, LB, LC, LD, BOYS, MODE, OP>), dim3((unsigned)g), dim3(64), 0, s, a);
    }
    return 0;
}

template <int LA, int LB, int LC, int LD, int OP>
int launch_coop_modes(int boys, int mode, const QuartetArgs& a, hipStream_t s)
{
    if (boys == 0) {
        if (mode == 0) return launch_coop<LA, LB, LC, LD, 0, 0, OP>(a, s);
        return launch_coop<LA, LB, LC, LD, 0, 1, OP>(a, s);
    }
    if (mode == 0) return launch_coop<LA, LB, LC, LD, 1, 0, OP>(a, s);
    return launch_coop<LA, LB, LC, LD, 1, 1, OP>(a, s);
}

// op: HFK_OP_JK / HFK_OP_GRAD (mode from the plan), HFK_OP_FLAT (screened
// when qtau > 0), HFK_OP_DIAG (unscreened).  FP64 only.
template <int LA, int LB, int LC, int LD>
int dispatch_coop(int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)
{
    if (prec != 0) return -1;
    switch (op) {
    case HFK_OP_JK: return launch_coop_modes<LA, LB, LC, LD, HFK_OP_JK>(boys, mode, a, s);
    case HFK_OP_GRAD: return launch_coop_modes<LA, LB, LC, LD, HFK_OP_GRAD>(boys, mode, a, s);
    case HFK_OP_FLAT: return launch_coop_modes<LA, LB, LC, LD, HFK_OP_FLAT>(boys, a.qtau > 0.0 ? 1 : 0, a, s);
    case HFK_OP_DIAG: return launch_coop_modes<LA, LB, LC, LD, HFK_OP_DIAG>(boys, 0, a, s);
    default: return -1;
    }
}

}  // namespace
