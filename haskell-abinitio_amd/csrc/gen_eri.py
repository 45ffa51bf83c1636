#!/usr/bin/env python3
"""Generate class-specialised McMurchie-Davidson ERI device code for gfx950.

For each angular-momentum class (la lb | lc ld) with la >= lb, lc >= ld and
(la,lb) >= (lc,ld) this emits

    template <typename T> __device__ void eri_<la><lb><lc><ld>(const EriArgs&, ..., T* V)

which evaluates one contracted shell quartet: it loops over the bra and ket
primitive pairs, evaluates the reference's Taylor-table Boys function (or the
accurate one), builds the Hermite integrals R_tuv by the recursion of
Integrals/IntegralsEvaluation.hs:430-449, contracts them with the ket and bra
Hermite expansion coefficients (IntegralsEvaluation.hs:452-477, same
mathematics as the reference's E.R.E product), applies the reference's
per-primitive, per-Cartesian-component screen (IntegralsEvaluation.hs:305-315)
and accumulates the contracted integrals V[(a*nb+b)*(nc*nd) + c*nd + d].

The generated code is fully unrolled straight-line arithmetic with every
index resolved at generation time, so the compiler keeps it in registers.

Usage: python gen_eri.py LMAX OUTDIR
"""
import os
import sys


def comps(l):
    """Cartesian components of angular momentum l in canonical order
    (x^l first, then decreasing x power, y before z): s; x,y,z;
    xx,xy,xz,yy,yz,zz."""
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


class Emitter:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self, expr, prefix="t"):
        name = "%s%d" % (prefix, self.n)
        self.n += 1
        self.lines.append("    const T %s = %s;" % (name, expr))
        return name

    def raw(self, s):
        self.lines.append(s)


def hermite_E(em, la, lb, PA, PB, oo2p, tag):
    """E[i][j][t] for i<=la, j<=lb (per axis); returns dict -> expr names.

    Helgaker's recursion (equivalent to recursiveHermCoeff with seed 1):
      E^{i+1,j}_t = oo2p*E^{ij}_{t-1} + PA*E^{ij}_t + (t+1)*E^{ij}_{t+1}
      E^{i,j+1}_t = oo2p*E^{ij}_{t-1} + PB*E^{ij}_t + (t+1)*E^{ij}_{t+1}
    """
    E = {(0, 0, 0): "T(1)"}

    def get(i, j, t):
        if t < 0 or t > i + j:
            return None
        return E[(i, j, t)]

    def build(i, j, src, X):
        # src = (i', j') with one fewer; X = PA or PB
        si, sj = src
        for t in range(0, i + j + 1):
            terms = []
            a = get(si, sj, t - 1)
            if a is not None:
                terms.append("%s*%s" % (oo2p, a) if a != "T(1)" else oo2p)
            b = get(si, sj, t)
            if b is not None:
                terms.append("%s*%s" % (X, b) if b != "T(1)" else X)
            c = get(si, sj, t + 1)
            if c is not None:
                terms.append("T(%d)*%s" % (t + 1, c) if t + 1 != 1 else c)
            expr = " + ".join(terms) if terms else "T(0)"
            E[(i, j, t)] = em.tmp(expr, "e%s" % tag)

    for i in range(1, la + 1):
        build(i, 0, (i - 1, 0), PA)
    for i in range(0, la + 1):
        for j in range(1, lb + 1):
            build(i, j, (i, j - 1), PB)
    return E


def gen_class(la, lb, lc, ld):
    L = la + lb + lc + ld
    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    na, nb, nc, nd = len(ca), len(cb), len(cc), len(cd)
    nbra, nket = na * nb, nc * nd
    name = "eri_%d%d%d%d" % (la, lb, lc, ld)
    out = []
    out.append("// (%s%s|%s%s): %d integrals, L=%d" % ("spdf"[la], "spdf"[lb], "spdf"[lc], "spdf"[ld], nbra * nket, L))
    out.append("// bra: nbp primitive-pair records (LDS or global); bq: their per-component")
    out.append("// screen values, bq[ib * %d + ab] (MODE 1)." % nbra)
    out.append("// kr: the ket's primitive-pair records (KetAoS or KetSoA).")
    out.append("// BRAK = 1: the bra records' PP_K field already holds 2 pi^(5/2) K.")
    out.append("template <typename T, int BOYS, int MODE, int BRAK, typename KR>")
    out.append("__device__ __forceinline__ void %s(const double* __restrict__ bra, const double* __restrict__ bq, int nbp," % name)
    out.append("    const KR& kr, int nkp, const BoysTables& bt, T* __restrict__ V, double qtau)")
    out.append("{")
    out.append("#pragma unroll")
    out.append("  for (int z = 0; z < %d; ++z) V[z] = T(0);" % (nbra * nket))
    # ket primitive pairs outer (per-lane data, loaded once each); bra
    # primitive pairs inner (wave-uniform: scalar loads)
    out.append("  for (int ik = 0; ik < nkp; ++ik) {")
    out.append("    const double* pk = kr.prim(ik);")
    em = Emitter()
    # a single-component ket uses QHI for QLO and for its component's screen
    # value: they equal it except that NaN is +inf there, and !(q_ab * v <
    # tau) holds for v = NaN and v = +inf alike (q_ab >= 0 or NaN)
    if nket > 1:
        em.raw("    const double kqlo = kr.at(pk, PP_QLO), kqhi = kr.at(pk, PP_QHI);")
    else:
        em.raw("    const double kqhi = kr.at(pk, PP_QHI), kqlo = kqhi;")
    em.raw("    const T kq = T(kr.at(pk, PP_P));")
    em.raw("    const T kQx = T(kr.at(pk, PP_PX)), kQy = T(kr.at(pk, PP_PY)), kQz = T(kr.at(pk, PP_PZ));")
    em.raw("    const T kCx = T(kr.at(pk, PP_AX)), kCy = T(kr.at(pk, PP_AY)), kCz = T(kr.at(pk, PP_AZ));")
    em.raw("    const T kDx = T(kr.at(pk, PP_BX)), kDy = T(kr.at(pk, PP_BY)), kDz = T(kr.at(pk, PP_BZ));")
    em.raw("    const T kK = T(kr.at(pk, PP_K));")
    em.raw("    const T koo2q = T(kr.at(pk, PP_OO2P));")
    if nket > 1 or nbra > 1:
        em.raw("    double kqv[%d];" % nket)
        em.raw("    if constexpr (MODE == 1) {")
        if nket > 1:
            em.raw("#pragma unroll")
            em.raw("      for (int z = 0; z < %d; ++z) kqv[z] = kr.q(pk, z);" % nket)
        else:
            em.raw("      kqv[0] = kqhi;")
        em.raw("    }")
    # the bra records are in descending order of QHI: a ket primitive screened
    # against the first one is screened against all, and so are the later ket
    # primitives (descending too)
    em.raw("    if (MODE == 1 && kqhi * bra[PP_QHI] < qtau) break;")
    Ek = [hermite_E(em, lc, ld, "kC" + ax, "kD" + ax, "koo2q", "k" + ax) for ax in "xyz"]
    out.extend(em.lines)
    def body(em2, sx, ibx):
        """One bra primitive pair (record pb<sx>, index ibx): returns
        {V index: (value name, bra component, ket component)}."""
        em2.raw("      const T bAx%s = T(pb%s[PP_AX]), bAy%s = T(pb%s[PP_AY]), bAz%s = T(pb%s[PP_AZ]);" % ((sx,) * 6))
        em2.raw("      const T bBx%s = T(pb%s[PP_BX]), bBy%s = T(pb%s[PP_BY]), bBz%s = T(pb%s[PP_BZ]);" % ((sx,) * 6))
        em2.raw("      const T bK%s = T(pb%s[PP_K]);" % (sx, sx))
        em2.raw("      const T boo2p%s = T(pb%s[PP_OO2P]);" % (sx, sx))
        if nket > 1 or nbra > 1:
            em2.raw("      double bqv%s[%d];" % (sx, nbra))
            em2.raw("      if constexpr (MODE == 1) {")
            em2.raw("#pragma unroll")
            em2.raw("        for (int z = 0; z < %d; ++z) bqv%s[z] = bq[(%s) * %d + z];" % (nbra, sx, ibx, nbra))
            em2.raw("      }")
        Eb = [hermite_E(em2, la, lb, "bA" + ax + sx, "bB" + ax + sx, "boo2p" + sx, "b" + ax) for ax in "xyz"]
        em2.raw("      T F%s[%d];" % (sx, L + 1))
        em2.raw("      boys_fast<T, %d, BOYS>(Targ%s, bt, F%s);" % (L, sx, sx))
        em2.raw("      const T pref%s = (BRAK ? bK%s : T(HFK_TWO_PI_2_5) * bK%s) * kK * rs%s;" % (sx, sx, sx, sx))
        # base R^n_000 = pref * (-2 alpha)^n * F_n
        em2.raw("      const T m2a%s = T(-2) * alpha%s;" % (sx, sx))
        R = {}
        pw = "pref" + sx
        for n in range(0, L + 1):
            if n == 0:
                R[(0, 0, 0, 0)] = em2.tmp("pref%s * F%s[0]" % (sx, sx), "r")
            else:
                pw = em2.tmp("%s * m2a%s" % (pw, sx), "w")
                R[(n, 0, 0, 0)] = em2.tmp("%s * F%s[%d]" % (pw, sx, n), "r")

        def getR(n, t, u, v):
            if t < 0 or u < 0 or v < 0:
                return None
            key = (n, t, u, v)
            if key in R:
                return R[key]
            if t >= 1:
                a_ = getR(n + 1, t - 2, u, v)
                b_ = getR(n + 1, t - 1, u, v)
                c_, X = t - 1, "Xpq"
            elif u >= 1:
                a_ = getR(n + 1, t, u - 2, v)
                b_ = getR(n + 1, t, u - 1, v)
                c_, X = u - 1, "Ypq"
            else:
                a_ = getR(n + 1, t, u, v - 2)
                b_ = getR(n + 1, t, u, v - 1)
                c_, X = v - 1, "Zpq"
            terms = []
            if a_ is not None and c_ != 0:
                terms.append("T(%d)*%s" % (c_, a_))
            terms.append("%s%s*%s" % (X, sx, b_))
            R[key] = em2.tmp(" + ".join(terms), "r")
            return R[key]

        lab = la + lb
        btuv = [(t, u, v) for t in range(lab + 1) for u in range(lab + 1 - t) for v in range(lab + 1 - t - u)]
        vals = {}
        for ic, c in enumerate(cc):
            for id_, d in enumerate(cd):
                ket_idx = ic * nd + id_
                # W^{cd}_{tuv} for tuv in btuv
                W = {}
                for (t, u, v) in btuv:
                    terms_x = []
                    for tau in range(c[0] + d[0] + 1):
                        ex = Ek[0][(c[0], d[0], tau)]
                        terms_y = []
                        for nu in range(c[1] + d[1] + 1):
                            ey = Ek[1][(c[1], d[1], nu)]
                            terms_z = []
                            for phi in range(c[2] + d[2] + 1):
                                ez = Ek[2][(c[2], d[2], phi)]
                                r = getR(0, t + tau, u + nu, v + phi)
                                sg = "-" if (phi % 2) else ""
                                terms_z.append("%s%s" % (sg, r) if ez == "T(1)" else "%s%s*%s" % (sg, ez, r))
                            sz = " + ".join(terms_z).replace("+ -", "- ")
                            sgy = (nu % 2) == 1
                            ty = "(%s)" % sz if ey == "T(1)" else "%s*(%s)" % (ey, sz)
                            terms_y.append(("-" if sgy else "") + ty)
                        sy = " + ".join(terms_y).replace("+ -", "- ")
                        sgx = (tau % 2) == 1
                        tx = "(%s)" % sy if ex == "T(1)" else "%s*(%s)" % (ex, sy)
                        terms_x.append(("-" if sgx else "") + tx)
                    sxp = " + ".join(terms_x).replace("+ -", "- ")
                    W[(t, u, v)] = em2.tmp(sxp, "W")
                for ia, a in enumerate(ca):
                    for ib_, b in enumerate(cb):
                        bra_idx = ia * nb + ib_
                        terms_x = []
                        for t in range(a[0] + b[0] + 1):
                            ex = Eb[0][(a[0], b[0], t)]
                            terms_y = []
                            for u in range(a[1] + b[1] + 1):
                                ey = Eb[1][(a[1], b[1], u)]
                                terms_z = []
                                for v in range(a[2] + b[2] + 1):
                                    ez = Eb[2][(a[2], b[2], v)]
                                    w = W[(t, u, v)]
                                    terms_z.append(w if ez == "T(1)" else "%s*%s" % (ez, w))
                                sz = " + ".join(terms_z)
                                terms_y.append("(%s)" % sz if ey == "T(1)" else "%s*(%s)" % (ey, sz))
                            sy = " + ".join(terms_y)
                            terms_x.append("(%s)" % sy if ex == "T(1)" else "%s*(%s)" % (ex, sy))
                        vals[bra_idx * nket + ket_idx] = (em2.tmp(" + ".join(terms_x), "v"), bra_idx, ket_idx)
        return vals

    em2 = Emitter()
    em2.n = em.n
    multi = nket > 1 or nbra > 1
    out.append("    for (int ib = 0; ib < nbp; ++ib) {")
    out.append("      const double* pb = bra + (size_t)ib * PP_STRIDE;")
    out.append("      const double bqlo = pb[PP_QLO], bqhi = pb[PP_QHI];")
    out.append("      // reference screen: component (ab|cd) of this primitive quartet is kept iff")
    out.append("      // !(q_ab*q_cd < tau).  MODE 0: no screen; MODE 1: the screen, branch-free.")
    out.append("      // A pair's records are in descending order of QHI, so once every component")
    out.append("      // is screened it stays screened for the rest of the bra primitives.")
    out.append("      if (MODE == 1 && bqhi * kqhi < qtau) break;")
    out.append("      T alpha, Xpq, Ypq, Zpq, rs, Targ;")
    out.append("      quartet_geom<T>(pb, kq, kQx, kQy, kQz, alpha, Xpq, Ypq, Zpq, rs, Targ);")
    out.append("      const bool allin = MODE == 0 || (bqlo * kqlo >= qtau);")
    vals = body(em2, "", "ib")
    if multi:
        # every active lane keeps every component: plain sums (a uniform
        # branch); otherwise the per-component selects
        em2.raw("      if (MODE == 0 || __all(allin)) {")
        for vi in sorted(vals):
            em2.raw("        V[%d] += %s;" % (vi, vals[vi][0]))
        em2.raw("      } else {")
        for vi in sorted(vals):
            name_, bi, ki = vals[vi]
            em2.raw("        V[%d] += (allin || !(bqv[%d] * kqv[%d] < qtau)) ? %s : T(0);" % (vi, bi, ki, name_))
        em2.raw("      }")
    else:
        for vi in sorted(vals):
            em2.raw("      V[%d] += %s;" % (vi, vals[vi][0]))
    out.extend(em2.lines)
    out.append("    }")
    out.append("  }")
    out.append("}")
    return "\n".join(out), (na, nb, nc, nd)


def prod_sum(factors_by_axis, inner):
    """sum over (t,u,v) of fx[t] fy[u] fz[v] inner(t,u,v), factorised per
    axis; factors are lists of (index, expr) ("T(1)" is a unit factor),
    inner returns an expression or None (no term).  None when empty."""
    tx = []
    for t, ex in factors_by_axis[0]:
        ty = []
        for u, ey in factors_by_axis[1]:
            tz = []
            for v, ez in factors_by_axis[2]:
                val = inner(t, u, v)
                if val is None:
                    continue
                tz.append(val if ez == "T(1)" else "%s*%s" % (ez, val))
            if tz:
                s = " + ".join(tz)
                ty.append("(%s)" % s if ey == "T(1)" else "%s*(%s)" % (ey, s))
        if ty:
            s = " + ".join(ty)
            tx.append("(%s)" % s if ex == "T(1)" else "%s*(%s)" % (ex, s))
    return " + ".join(tx).replace("+ -", "- ") if tx else None


# Classes whose gradient code forms R, W and X inside each branch of the
# all-components test (shorter live ranges; measured faster for these), the
# others form them once before the branch.
GRAD_LAZY = {(1, 1, 1, 0)}
# Classes whose gradient splits the bra components over this many launches
# (gen_grad_class_z): the density block of all components does not fit one
# lane's registers.
GRAD_SPLIT = {(1, 1, 1, 1): 3}


def gen_grad_class(la, lb, lc, ld):
    """Derivative-integral contraction for the nuclear gradient of class
    (la lb|lc ld) over one bra pair and one ket pair, with the two-particle
    density G[ab][cd] of the shell quartet.  Per primitive quartet
    (McMurchie-Davidson, bra Omega_ab = K sum_t E^ab_t Lambda_t(P)):

      dA:   2 alpha g_{a+1} - a g_{a-1}   (Hermite coefficients
            DA^{ab}_t = 2 alpha E^{a+1,b}_t - a E^{a-1,b}_t, t <= a+b+1)
      dA + dB = dP: E^ab_tuv with R shifted by one along the axis
      dC = (gamma/q) dQ + d/dX_CD, dQ = -dP, and d/dX_CD of the ket's
            coefficients dE^cd_t = 2 gamma E^{c+1,d}_t - c E^{c-1,d}_t
            - (gamma/q) E^{cd}_{t-1} (t <= c+d)
      dD = -(dA + dB + dC)   (translational invariance; formed by the caller)

    The ket tables depend on the ket primitive only, so the d/dX_CD part is
    applied once per ket primitive to the bra-loop sums of
    Z^cd_tuv = sum_ab G X^ab_tuv.  Outputs g[0..2] = dA, g[3..5] = dB,
    g[6..8] = dC (each contracted with G, summed over the primitive
    quartets).  Spec of the derivative: Derivatives.hs:246-257."""
    L = la + lb + lc + ld
    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    na, nb, nc, nd = len(ca), len(cb), len(cc), len(cd)
    nbra, nket = na * nb, nc * nd
    lab, lcd = la + lb, lc + ld
    multi = nket > 1 or nbra > 1
    name = "grad_%d%d%d%d" % (la, lb, lc, ld)
    out = []
    out.append("// bra: nbp staged primitive-pair records (PP_K holds 2 pi^(5/2) K); bq: their")
    out.append("// per-component screen values, bq[ib * %d + ab] (MODE 1)." % nbra)
    out.append("template <typename T, int BOYS, int MODE, typename KR>")
    out.append("__device__ __forceinline__ void %s(const double* __restrict__ bra, const double* __restrict__ bq, int nbp," % name)
    out.append("    const KR& kr, int nkp, const BoysTables& bt, const T* __restrict__ G, double qtau, T* __restrict__ g)")
    out.append("{")
    out.append("  T gA0 = T(0), gA1 = T(0), gA2 = T(0), gP0 = T(0), gP1 = T(0), gP2 = T(0);")
    out.append("  T gC0 = T(0), gC1 = T(0), gC2 = T(0);")
    out.append("  for (int ik = 0; ik < nkp; ++ik) {")
    out.append("    const double* pk = kr.prim(ik);")
    em = Emitter()
    if nket > 1:
        em.raw("    const double kqlo = kr.at(pk, PP_QLO), kqhi = kr.at(pk, PP_QHI);")
    else:
        em.raw("    const double kqhi = kr.at(pk, PP_QHI), kqlo = kqhi;")
    em.raw("    const T kq = T(kr.at(pk, PP_P));")
    em.raw("    const T kQx = T(kr.at(pk, PP_PX)), kQy = T(kr.at(pk, PP_PY)), kQz = T(kr.at(pk, PP_PZ));")
    em.raw("    const T kCx = T(kr.at(pk, PP_AX)), kCy = T(kr.at(pk, PP_AY)), kCz = T(kr.at(pk, PP_AZ));")
    em.raw("    const T kDx = T(kr.at(pk, PP_BX)), kDy = T(kr.at(pk, PP_BY)), kDz = T(kr.at(pk, PP_BZ));")
    em.raw("    const T kK = T(kr.at(pk, PP_K));")
    em.raw("    const T koo2q = T(kr.at(pk, PP_OO2P));")
    em.raw("    const T kgam2 = T(2) * T(kr.at(pk, PP_EA));")
    em.raw("    const T kgoq = kgam2 * koo2q;   // gamma / q")
    if multi:
        em.raw("    double kqv[%d];" % nket)
        em.raw("    if constexpr (MODE == 1) {")
        if nket > 1:
            em.raw("#pragma unroll")
            em.raw("      for (int z = 0; z < %d; ++z) kqv[z] = kr.q(pk, z);" % nket)
        else:
            em.raw("      kqv[0] = kqhi;")
        em.raw("    }")
    em.raw("    if (MODE == 1 && kqhi * bra[PP_QHI] < qtau) break;   // records by descending QHI")
    Ek = [hermite_E(em, lc + 1, ld, "kC" + ax, "kD" + ax, "koo2q", "k" + ax) for ax in "xyz"]

    def Efac(E, i, j):
        return [(t, E[(i, j, t)]) for t in range(i + j + 1)]

    # ket d/dX_CD tables, per axis and (c, d) exponent pair
    dEk = [dict() for _ in range(3)]
    for x in range(3):
        for c in cc:
            for d in cd:
                key = (c[x], d[x])
                if key in dEk[x]:
                    continue
                ci, di = key
                lst = []
                for t in range(ci + di + 1):
                    terms = []
                    e = Ek[x][(ci + 1, di, t)]
                    terms.append("kgam2*%s" % e if e != "T(1)" else "kgam2")
                    if ci > 0 and t <= ci - 1 + di:
                        e = Ek[x][(ci - 1, di, t)]
                        terms.append("- T(%d)*%s" % (ci, e) if e != "T(1)" else "- T(%d)" % ci)
                    if t >= 1:
                        e = Ek[x][(ci, di, t - 1)]
                        terms.append("- kgoq*%s" % e if e != "T(1)" else "- kgoq")
                    lst.append((t, em.tmp(" ".join(terms), "dk")))
                dEk[x][key] = lst
    # bra-loop sums of Z^cd_tuv (tuv within the ket's Hermite box)
    zbox = {}
    for ic, c in enumerate(cc):
        for id_, d in enumerate(cd):
            kidx = ic * nd + id_
            zbox[kidx] = [(t, u, v) for t in range(c[0] + d[0] + 1) for u in range(c[1] + d[1] + 1)
                          for v in range(c[2] + d[2] + 1)]
            for (t, u, v) in zbox[kidx]:
                em.raw("    T za%d_%d%d%d = T(0);" % (kidx, t, u, v))
    em.raw("    T sp0 = T(0), sp1 = T(0), sp2 = T(0);")
    out.extend(em.lines)

    out.append("    for (int ib = 0; ib < nbp; ++ib) {")
    out.append("      const double* pb = bra + (size_t)ib * PP_STRIDE;")
    out.append("      const double bqlo = pb[PP_QLO], bqhi = pb[PP_QHI];")
    out.append("      if (MODE == 1 && bqhi * kqhi < qtau) break;   // records by descending QHI")
    out.append("      T alpha, Xpq, Ypq, Zpq, rs, Targ;")
    out.append("      quartet_geom<T>(pb, kq, kQx, kQy, kQz, alpha, Xpq, Ypq, Zpq, rs, Targ);")
    out.append("      const bool allin = MODE == 0 || (bqlo * kqlo >= qtau);")
    em2 = Emitter()
    em2.n = em.n
    em2.raw("      const T bAx = T(pb[PP_AX]), bAy = T(pb[PP_AY]), bAz = T(pb[PP_AZ]);")
    em2.raw("      const T bBx = T(pb[PP_BX]), bBy = T(pb[PP_BY]), bBz = T(pb[PP_BZ]);")
    em2.raw("      const T bK = T(pb[PP_K]);")
    em2.raw("      const T boo2p = T(pb[PP_OO2P]);")
    em2.raw("      const T balp2 = T(2) * T(pb[PP_EA]);")
    Eb = [hermite_E(em2, la + 1, lb, "bA" + ax, "bB" + ax, "boo2p", "b" + ax) for ax in "xyz"]
    em2.raw("      T F[%d];" % (L + 2))
    em2.raw("      boys_fast<T, %d, BOYS>(Targ, bt, F);" % (L + 1))
    em2.raw("      const T pref = bK * kK * rs;")
    em2.raw("      const T m2a = T(-2) * alpha;")
    R = {}
    pw = "pref"
    for n in range(0, L + 2):
        if n == 0:
            R[(0, 0, 0, 0)] = em2.tmp("pref * F[0]", "r")
        else:
            pw = em2.tmp("%s * m2a" % pw, "w")
            R[(n, 0, 0, 0)] = em2.tmp("%s * F[%d]" % (pw, n), "r")

    def getR(Rm, n, t, u, v):
        """R^n_tuv, memoised in Rm (one memo per branch: a value formed
        inside one branch is not visible in the other)."""
        if t < 0 or u < 0 or v < 0:
            return None
        key = (n, t, u, v)
        if key in Rm:
            return Rm[key]
        if t >= 1:
            a_ = getR(Rm, n + 1, t - 2, u, v); b_ = getR(Rm, n + 1, t - 1, u, v); X = "Xpq"; c_ = t - 1
        elif u >= 1:
            a_ = getR(Rm, n + 1, t, u - 2, v); b_ = getR(Rm, n + 1, t, u - 1, v); X = "Ypq"; c_ = u - 1
        else:
            a_ = getR(Rm, n + 1, t, u, v - 2); b_ = getR(Rm, n + 1, t, u, v - 1); X = "Zpq"; c_ = v - 1
        terms = []
        if a_ is not None and c_ != 0:
            terms.append("T(%d)*%s" % (c_, a_))
        terms.append("%s*%s" % (X, b_))
        Rm[key] = em2.tmp(" + ".join(terms), "r")
        return Rm[key]

    # W^{cd}_{tuv} = sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}
    def getW(Rm, Wc, kidx, t, u, v):
        key = (kidx, t, u, v)
        if key in Wc:
            return Wc[key]
        c, d = cc[kidx // nd], cd[kidx % nd]
        fac = [Efac(Ek[x], c[x], d[x]) for x in range(3)]

        def inner(tau, nu, phi):
            r = getR(Rm, 0, t + tau, u + nu, v + phi)
            return ("-%s" % r) if (tau + nu + phi) % 2 else r
        Wc[key] = em2.tmp(prod_sum(fac, inner), "W")
        return Wc[key]

    # bra dA tables per axis and (a, b) exponent pair: t <= a+b+1
    DA = [dict() for _ in range(3)]
    for x in range(3):
        for a in ca:
            for b in cb:
                key = (a[x], b[x])
                if key in DA[x]:
                    continue
                ai, bi = key
                lst = []
                for t in range(ai + bi + 2):
                    e = Eb[x][(ai + 1, bi, t)]
                    terms = ["balp2*%s" % e if e != "T(1)" else "balp2"]
                    if ai > 0 and t <= ai - 1 + bi:
                        e = Eb[x][(ai - 1, bi, t)]
                        terms.append("- T(%d)*%s" % (ai, e) if e != "T(1)" else "- T(%d)" % ai)
                    lst.append((t, em2.tmp(" ".join(terms), "da")))
                DA[x][key] = lst
    # X^{ab}_{tau nu phi} = sum_tuv E^{ab}_tuv R_{t+tau,...}, over the union of the ket boxes
    kunion = sorted(set(k for b_ in zbox.values() for k in b_))

    def getX(Rm, Xv, bidx, tau, nu, phi):
        key = (bidx, tau, nu, phi)
        if key not in Xv:
            a, b = ca[bidx // nb], cb[bidx % nb]
            fac = [Efac(Eb[x], a[x], b[x]) for x in range(3)]
            Xv[key] = em2.tmp(prod_sum(fac, lambda t, u, v: getR(Rm, 0, t + tau, u + nu, v + phi)), "X")
        return Xv[key]
    # Y index sets; W and X are formed where first used inside each branch,
    # which keeps their live ranges short
    ysets = {}
    for ia, a in enumerate(ca):
        for ib_, b in enumerate(cb):
            bidx = ia * nb + ib_
            box = [a[x] + b[x] for x in range(3)]
            s = []
            for t in range(box[0] + 2):
                for u in range(box[1] + 2):
                    for v in range(box[2] + 2):
                        over = max(0, t - box[0]) + max(0, u - box[1]) + max(0, v - box[2])
                        if over <= 1:
                            s.append((t, u, v))
            ysets[bidx] = s

    lazy = (la, lb, lc, ld) in GRAD_LAZY
    shared = (R, {}, {})
    if not lazy:
        for bidx in range(nbra):
            for (t, u, v) in ysets[bidx]:
                for k in range(nket):
                    getW(R, shared[1], k, t, u, v)
            for (t, u, v) in kunion:
                getX(R, shared[2], bidx, t, u, v)

    def body(Gname, ind):
        """Y, the dA and dP sums and the Z sums with density names Gname(ab, cd)."""
        Rm, Wc, Xv = (dict(R), {}, {}) if lazy else shared
        for ia, a in enumerate(ca):
            for ib_, b in enumerate(cb):
                bidx = ia * nb + ib_
                Y = {}
                for (t, u, v) in ysets[bidx]:
                    terms = ["%s*%s" % (Gname(bidx, k), getW(Rm, Wc, k, t, u, v)) for k in range(nket)]
                    Y[(t, u, v)] = em2.tmp(" + ".join(terms), "Y")
                for k in range(3):
                    fac = [DA[x][(a[x], b[x])] if x == k else Efac(Eb[x], a[x], b[x]) for x in range(3)]
                    val = prod_sum(fac, lambda t, u, v: Y.get((t, u, v)))
                    if val:
                        em2.raw("%sgA%d += %s;" % (ind, k, val))
                    fac = [Efac(Eb[x], a[x], b[x]) for x in range(3)]
                    sh = [1 if x == k else 0 for x in range(3)]
                    val = prod_sum(fac, lambda t, u, v: Y.get((t + sh[0], u + sh[1], v + sh[2])))
                    if val:
                        em2.raw("%ssp%d += %s;" % (ind, k, val))
        for kidx in range(nket):
            for (t, u, v) in zbox[kidx]:
                terms = ["%s*%s" % (Gname(bb, kidx), getX(Rm, Xv, bb, t, u, v)) for bb in range(nbra)]
                em2.raw("%sza%d_%d%d%d += %s;" % (ind, kidx, t, u, v, " + ".join(terms)))

    if multi:
        em2.raw("      double bqv[%d];" % nbra)
        em2.raw("      if constexpr (MODE == 1) {")
        em2.raw("#pragma unroll")
        em2.raw("        for (int z = 0; z < %d; ++z) bqv[z] = bq[ib * %d + z];" % (nbra, nbra))
        em2.raw("      }")
        # every active lane keeps every component: plain density (a uniform
        # branch); otherwise the per-component selects
        em2.raw("      if (MODE == 0 || __all(allin)) {")
        body(lambda ab, k: "T(G[%d])" % (ab * nket + k), "        ")
        em2.raw("      } else {")
        Gm = {}
        for ab in range(nbra):
            for k in range(nket):
                Gm[(ab, k)] = em2.tmp("(allin || !(bqv[%d] * kqv[%d] < qtau)) ? T(G[%d]) : T(0)" % (ab, k, ab * nket + k), "gm")
        body(lambda ab, k: Gm[(ab, k)], "        ")
        em2.raw("      }")
    else:
        body(lambda ab, k: "T(G[0])", "      ")
    out.extend(em2.lines)
    out.append("    }")
    # per ket primitive: dP and the d/dX_CD part of dC
    em3 = Emitter()
    em3.n = em2.n
    for k in range(3):
        terms = []
        for kidx in range(nket):
            c, d = cc[kidx // nd], cd[kidx % nd]
            fac = [dEk[x][(c[x], d[x])] if x == k else Efac(Ek[x], c[x], d[x]) for x in range(3)]

            def inner(t, u, v, kidx=kidx):
                z = "za%d_%d%d%d" % (kidx, t, u, v)
                return ("-%s" % z) if (t + u + v) % 2 else z
            val = prod_sum(fac, inner)
            if val:
                terms.append(val)
        scd = " + ".join(terms).replace("+ -", "- ") if terms else "T(0)"
        em3.raw("    gP%d += sp%d;" % (k, k))
        em3.raw("    gC%d += (%s) - kgoq * sp%d;" % (k, scd, k))
    out.extend(em3.lines)
    out.append("  }")
    out.append("  g[0] = gA0; g[1] = gA1; g[2] = gA2;")
    out.append("  g[3] = gP0 - gA0; g[4] = gP1 - gA1; g[5] = gP2 - gA2;")
    out.append("  g[6] = gC0; g[7] = gC1; g[8] = gC2;")
    out.append("}")
    return "\n".join(out)


def gen_grad_class_z(la, lb, lc, ld, sub, fname):
    """Gradient contraction for the bra components `sub` (indices into the
    bra's Cartesian pairs) of class (la lb|lc ld), for classes whose ket
    holds too many components for the W route.  G is the lane's density
    block G[s][cd] over s in sub.  Per ket primitive the density is
    contracted with the ket's Hermite coefficients,
        Z^s_tuv = sum_cd G[s][cd] (-1)^{t+u+v} E^cd_tuv,
    so per primitive quartet Y^s_tuv = sum_tau Z^s_tau R_{t+tau} (no W);
    dA, dP and the Z^cd sums for dC follow gen_grad_class.  The outputs are
    the sub-block's share of dA, dB, dC: partial sums over bra components,
    which the caller adds."""
    L = la + lb + lc + ld
    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    na, nb, nc, nd = len(ca), len(cb), len(cc), len(cd)
    nbra, nket = na * nb, nc * nd
    nsub = len(sub)
    lab, lcd = la + lb, lc + ld
    out = []
    out.append("// bra: nbp staged primitive-pair records (PP_K holds 2 pi^(5/2) K); bq: their")
    out.append("// per-component screen values, bq[ib * %d + ab] (MODE 1).  G: [%d][%d], bra" % (nbra, nsub, nket))
    out.append("// components %s." % ", ".join(str(x) for x in sub))
    out.append("template <typename T, int BOYS, int MODE, typename KR>")
    out.append("__device__ __forceinline__ void %s(const double* __restrict__ bra, const double* __restrict__ bq, int nbp," % fname)
    out.append("    const KR& kr, int nkp, const BoysTables& bt, const T* __restrict__ G, double qtau, T* __restrict__ g)")
    out.append("{")
    out.append("  T gA0 = T(0), gA1 = T(0), gA2 = T(0), gP0 = T(0), gP1 = T(0), gP2 = T(0);")
    out.append("  T gC0 = T(0), gC1 = T(0), gC2 = T(0);")
    out.append("  for (int ik = 0; ik < nkp; ++ik) {")
    out.append("    const double* pk = kr.prim(ik);")
    em = Emitter()
    em.raw("    const double kqlo = kr.at(pk, PP_QLO), kqhi = kr.at(pk, PP_QHI);")
    em.raw("    const T kq = T(kr.at(pk, PP_P));")
    em.raw("    const T kQx = T(kr.at(pk, PP_PX)), kQy = T(kr.at(pk, PP_PY)), kQz = T(kr.at(pk, PP_PZ));")
    em.raw("    const T kCx = T(kr.at(pk, PP_AX)), kCy = T(kr.at(pk, PP_AY)), kCz = T(kr.at(pk, PP_AZ));")
    em.raw("    const T kDx = T(kr.at(pk, PP_BX)), kDy = T(kr.at(pk, PP_BY)), kDz = T(kr.at(pk, PP_BZ));")
    em.raw("    const T kK = T(kr.at(pk, PP_K));")
    em.raw("    const T koo2q = T(kr.at(pk, PP_OO2P));")
    em.raw("    const T kgam2 = T(2) * T(kr.at(pk, PP_EA));")
    em.raw("    const T kgoq = kgam2 * koo2q;   // gamma / q")
    em.raw("    double kqv[%d];" % nket)
    em.raw("    if constexpr (MODE == 1) {")
    em.raw("#pragma unroll")
    em.raw("      for (int z = 0; z < %d; ++z) kqv[z] = kr.q(pk, z);" % nket)
    em.raw("    }")
    em.raw("    if (MODE == 1 && kqhi * bra[PP_QHI] < qtau) break;   // records by descending QHI")
    Ek = [hermite_E(em, lc, ld, "kC" + ax, "kD" + ax, "koo2q", "k" + ax) for ax in "xyz"]

    def Efac(E, i, j):
        return [(t, E[(i, j, t)]) for t in range(i + j + 1)]

    zbox = {}
    for ic, c in enumerate(cc):
        for id_, d in enumerate(cd):
            kidx = ic * nd + id_
            zbox[kidx] = [(t, u, v) for t in range(c[0] + d[0] + 1) for u in range(c[1] + d[1] + 1)
                          for v in range(c[2] + d[2] + 1)]
    kunion = sorted(set(k for b_ in zbox.values() for k in b_))
    # ket products (-1)^{t+u+v} E^cd_tuv, once per ket primitive
    Pk = {}
    for kidx in range(nket):
        c, d = cc[kidx // nd], cd[kidx % nd]
        for (t, u, v) in zbox[kidx]:
            fs = [f for f in (Ek[0][(c[0], d[0], t)], Ek[1][(c[1], d[1], u)], Ek[2][(c[2], d[2], v)]) if f != "T(1)"]
            e = "*".join(fs) if fs else "T(1)"
            if (t + u + v) % 2:
                e = "-(%s)" % e if e != "T(1)" else "T(-1)"
            Pk[(kidx, t, u, v)] = em.tmp(e, "pk") if ("*" in e or e.startswith("-")) else e

    def zsum(emx, Gname, tag):
        Z = {}
        for si in range(nsub):
            for key in kunion:
                terms = []
                for kidx in range(nket):
                    if (kidx,) + key in Pk:
                        p = Pk[(kidx,) + key]
                        terms.append(Gname(si, kidx) if p == "T(1)" else "%s*%s" % (Gname(si, kidx), p))
                Z[(si,) + key] = emx.tmp(" + ".join(terms), tag)
        return Z
    Zf = zsum(em, lambda si, k: "T(G[%d])" % (si * nket + k), "zf")
    for kidx in range(nket):
        for (t, u, v) in zbox[kidx]:
            em.raw("    T za%d_%d%d%d = T(0);" % (kidx, t, u, v))
    em.raw("    T sp0 = T(0), sp1 = T(0), sp2 = T(0);")
    out.extend(em.lines)

    out.append("    for (int ib = 0; ib < nbp; ++ib) {")
    out.append("      const double* pb = bra + (size_t)ib * PP_STRIDE;")
    out.append("      const double bqlo = pb[PP_QLO], bqhi = pb[PP_QHI];")
    out.append("      if (MODE == 1 && bqhi * kqhi < qtau) break;   // records by descending QHI")
    out.append("      T alpha, Xpq, Ypq, Zpq, rs, Targ;")
    out.append("      quartet_geom<T>(pb, kq, kQx, kQy, kQz, alpha, Xpq, Ypq, Zpq, rs, Targ);")
    out.append("      const bool allin = MODE == 0 || (bqlo * kqlo >= qtau);")
    em2 = Emitter()
    em2.n = em.n
    em2.raw("      const T bAx = T(pb[PP_AX]), bAy = T(pb[PP_AY]), bAz = T(pb[PP_AZ]);")
    em2.raw("      const T bBx = T(pb[PP_BX]), bBy = T(pb[PP_BY]), bBz = T(pb[PP_BZ]);")
    em2.raw("      const T bK = T(pb[PP_K]);")
    em2.raw("      const T boo2p = T(pb[PP_OO2P]);")
    em2.raw("      const T balp2 = T(2) * T(pb[PP_EA]);")
    Eb = [hermite_E(em2, la + 1, lb, "bA" + ax, "bB" + ax, "boo2p", "b" + ax) for ax in "xyz"]
    em2.raw("      T F[%d];" % (L + 2))
    em2.raw("      boys_fast<T, %d, BOYS>(Targ, bt, F);" % (L + 1))
    em2.raw("      const T pref = bK * kK * rs;")
    em2.raw("      const T m2a = T(-2) * alpha;")
    R = {}
    pw = "pref"
    for n in range(0, L + 2):
        if n == 0:
            R[(0, 0, 0, 0)] = em2.tmp("pref * F[0]", "r")
        else:
            pw = em2.tmp("%s * m2a" % pw, "w")
            R[(n, 0, 0, 0)] = em2.tmp("%s * F[%d]" % (pw, n), "r")

    def getR(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return None
        key = (n, t, u, v)
        if key in R:
            return R[key]
        if t >= 1:
            a_ = getR(n + 1, t - 2, u, v); b_ = getR(n + 1, t - 1, u, v); X = "Xpq"; c_ = t - 1
        elif u >= 1:
            a_ = getR(n + 1, t, u - 2, v); b_ = getR(n + 1, t, u - 1, v); X = "Ypq"; c_ = u - 1
        else:
            a_ = getR(n + 1, t, u, v - 2); b_ = getR(n + 1, t, u, v - 1); X = "Zpq"; c_ = v - 1
        terms = []
        if a_ is not None and c_ != 0:
            terms.append("T(%d)*%s" % (c_, a_))
        terms.append("%s*%s" % (X, b_))
        R[key] = em2.tmp(" + ".join(terms), "r")
        return R[key]

    DA = [dict() for _ in range(3)]
    for x in range(3):
        for s in sub:
            a, b = ca[s // nb], cb[s % nb]
            key = (a[x], b[x])
            if key in DA[x]:
                continue
            ai, bi = key
            lst = []
            for t in range(ai + bi + 2):
                e = Eb[x][(ai + 1, bi, t)]
                terms = ["balp2*%s" % e if e != "T(1)" else "balp2"]
                if ai > 0 and t <= ai - 1 + bi:
                    e = Eb[x][(ai - 1, bi, t)]
                    terms.append("- T(%d)*%s" % (ai, e) if e != "T(1)" else "- T(%d)" % ai)
                lst.append((t, em2.tmp(" ".join(terms), "da")))
            DA[x][key] = lst
    ysets = {}
    for si, s in enumerate(sub):
        a, b = ca[s // nb], cb[s % nb]
        box = [a[x] + b[x] for x in range(3)]
        ys = []
        for t in range(box[0] + 2):
            for u in range(box[1] + 2):
                for v in range(box[2] + 2):
                    over = max(0, t - box[0]) + max(0, u - box[1]) + max(0, v - box[2])
                    if over <= 1:
                        ys.append((t, u, v))
        ysets[si] = ys
    # R values used below, formed before the branch
    for si in range(nsub):
        for (t, u, v) in ysets[si]:
            for (tau, nu, phi) in kunion:
                getR(0, t + tau, u + nu, v + phi)
    Xv = {}
    for si, s in enumerate(sub):
        a, b = ca[s // nb], cb[s % nb]
        fac = [Efac(Eb[x], a[x], b[x]) for x in range(3)]
        for (tau, nu, phi) in kunion:
            Xv[(si, tau, nu, phi)] = em2.tmp(
                prod_sum(fac, lambda t, u, v, tau=tau, nu=nu, phi=phi: getR(0, t + tau, u + nu, v + phi)), "X")

    def body(Z, Gname, ind):
        for si, s in enumerate(sub):
            a, b = ca[s // nb], cb[s % nb]
            Y = {}
            for (t, u, v) in ysets[si]:
                terms = []
                for (tau, nu, phi) in kunion:
                    r = getR(0, t + tau, u + nu, v + phi)
                    terms.append("%s*%s" % (Z[(si, tau, nu, phi)], r))
                Y[(t, u, v)] = em2.tmp(" + ".join(terms), "Y")
            for k in range(3):
                fac = [DA[x][(a[x], b[x])] if x == k else Efac(Eb[x], a[x], b[x]) for x in range(3)]
                val = prod_sum(fac, lambda t, u, v: Y.get((t, u, v)))
                if val:
                    em2.raw("%sgA%d += %s;" % (ind, k, val))
                fac = [Efac(Eb[x], a[x], b[x]) for x in range(3)]
                sh = [1 if x == k else 0 for x in range(3)]
                val = prod_sum(fac, lambda t, u, v: Y.get((t + sh[0], u + sh[1], v + sh[2])))
                if val:
                    em2.raw("%ssp%d += %s;" % (ind, k, val))
        for kidx in range(nket):
            for (t, u, v) in zbox[kidx]:
                terms = ["%s*%s" % (Gname(si, kidx), Xv[(si, t, u, v)]) for si in range(nsub)]
                em2.raw("%sza%d_%d%d%d += %s;" % (ind, kidx, t, u, v, " + ".join(terms)))

    em2.raw("      double bqv[%d];" % nsub)
    em2.raw("      if constexpr (MODE == 1) {")
    for si, s in enumerate(sub):
        em2.raw("        bqv[%d] = bq[ib * %d + %d];" % (si, nbra, s))
    em2.raw("      }")
    em2.raw("      if (MODE == 0 || __all(allin)) {")
    body(Zf, lambda si, k: "T(G[%d])" % (si * nket + k), "        ")
    em2.raw("      } else {")
    gm = {}
    for si in range(nsub):
        for k in range(nket):
            gm[(si, k)] = em2.tmp("(allin || !(bqv[%d] * kqv[%d] < qtau)) ? T(G[%d]) : T(0)" % (si, k, si * nket + k), "gm")
    Zm = zsum(em2, lambda si, k: gm[(si, k)], "zm")
    body(Zm, lambda si, k: gm[(si, k)], "        ")
    em2.raw("      }")
    out.extend(em2.lines)
    out.append("    }")
    # per ket primitive: dP and the d/dX_CD part of dC (ket tables formed again here)
    em3 = Emitter()
    em3.n = em2.n
    Ek2 = [hermite_E(em3, lc + 1, ld, "kC" + ax, "kD" + ax, "koo2q", "q" + ax) for ax in "xyz"]
    dEk = [dict() for _ in range(3)]
    for x in range(3):
        for c in cc:
            for d in cd:
                key = (c[x], d[x])
                if key in dEk[x]:
                    continue
                ci, di = key
                lst = []
                for t in range(ci + di + 1):
                    terms = []
                    e = Ek2[x][(ci + 1, di, t)]
                    terms.append("kgam2*%s" % e if e != "T(1)" else "kgam2")
                    if ci > 0 and t <= ci - 1 + di:
                        e = Ek2[x][(ci - 1, di, t)]
                        terms.append("- T(%d)*%s" % (ci, e) if e != "T(1)" else "- T(%d)" % ci)
                    if t >= 1:
                        e = Ek2[x][(ci, di, t - 1)]
                        terms.append("- kgoq*%s" % e if e != "T(1)" else "- kgoq")
                    lst.append((t, em3.tmp(" ".join(terms), "dk")))
                dEk[x][key] = lst
    for k in range(3):
        terms = []
        for kidx in range(nket):
            c, d = cc[kidx // nd], cd[kidx % nd]
            fac = [dEk[x][(c[x], d[x])] if x == k else Efac(Ek2[x], c[x], d[x]) for x in range(3)]

            def inner(t, u, v, kidx=kidx):
                z = "za%d_%d%d%d" % (kidx, t, u, v)
                return ("-%s" % z) if (t + u + v) % 2 else z
            val = prod_sum(fac, inner)
            if val:
                terms.append(val)
        scd = " + ".join(terms).replace("+ -", "- ") if terms else "T(0)"
        em3.raw("    gP%d += sp%d;" % (k, k))
        em3.raw("    gC%d += (%s) - kgoq * sp%d;" % (k, scd, k))
    out.extend(em3.lines)
    out.append("  }")
    out.append("  g[0] = gA0; g[1] = gA1; g[2] = gA2;")
    out.append("  g[3] = gP0 - gA0; g[4] = gP1 - gA1; g[5] = gP2 - gA2;")
    out.append("  g[6] = gC0; g[7] = gC1; g[8] = gC2;")
    out.append("}")
    return "\n".join(out)


# Classes with a d shell holding more contracted integrals per quartet than
# this run on the cooperative wave kernel (per-lane registers cannot hold
# them); the host's hfk_class_coop() uses the same bounds.
COOP_NV_JK = 54
COOP_NV_GRAD = 18


def pair_classes(lmax):
    return [(a, b) for a in range(lmax + 1) for b in range(a + 1)]


def quartet_classes(lmax):
    pcs = pair_classes(lmax)
    out = []
    for i, (la, lb) in enumerate(pcs):
        for (lc, ld) in pcs[: i + 1]:
            out.append((la, lb, lc, ld))
    return out


def main():
    """gen_eri.py LMAX OUTDIR: writes OUTDIR/eri_<class>.hip (one translation
    unit per class, compiled in parallel) and OUTDIR/eri_dispatch.hip."""
    lmax = int(sys.argv[1]) if len(sys.argv) > 1 else 1
    outdir = sys.argv[2] if len(sys.argv) > 2 else "."
    import os
    os.makedirs(outdir, exist_ok=True)
    qs = quartet_classes(lmax)
    for q in qs:
        tag = "%d%d%d%d" % q
        nv = 1
        for l in q:
            nv *= (l + 1) * (l + 2) // 2
        jk_coop = max(q) >= 2 and nv > COOP_NV_JK   # cooperative wave kernel (hfk_coop_impl.hpp)
        grad_coop = max(q) >= 2 and nv > COOP_NV_GRAD
        body = ["// Generated by gen_eri.py -- do not edit.", "#include <hip/hip_runtime.h>"]
        if jk_coop:
            body += ['#include "hfk_coop_impl.hpp"',
                     "int hfk_dispatch_%s(int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)" % tag,
                     "{ return dispatch_coop<%d, %d, %d, %d>(op, prec, boys, mode, a, s); }" % q,
                     "int hfk_dispatch_grad_%s(int boys, int mode, int, const QuartetArgs& a, hipStream_t s)" % tag,
                     "{ return dispatch_coop<%d, %d, %d, %d>(HFK_OP_GRAD, 0, boys, mode, a, s); }" % q]
            path = os.path.join(outdir, "eri_%s.hip" % tag)
            txt = "\n".join(body) + "\n"
            if not os.path.exists(path) or open(path).read() != txt:
                with open(path, "w") as f:
                    f.write(txt)
            continue
        src, _ = gen_class(*q)
        body.append('#include "hfk_internal.hpp"')
        body.append("namespace {")
        body.append(src)
        body.append("template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE, int BRAK, typename KR>")
        body.append("__device__ __forceinline__ void eri_class(const double* __restrict__ bra, const double* __restrict__ bq, int nbp,")
        body.append("    const KR& kr, int nkp, const BoysTables& bt, T* __restrict__ V, double qtau)")
        body.append("{ eri_%s<T, BOYS, MODE, BRAK>(bra, bq, nbp, kr, nkp, bt, V, qtau); }" % tag)
        if not grad_coop:
            ngrp = GRAD_SPLIT.get(q, 1)
            body.append("// bra components per gradient launch: NBRA / kGradGroups")
            body.append("constexpr int kGradGroups = %d;" % ngrp)
            if ngrp == 1:
                body.append(gen_grad_class(*q))
                calls = ["grad_%s<T, BOYS, MODE>(bra, bq, nbp, kr, nkp, bt, G, qtau, g);" % tag]
            else:
                na_ = (q[0] + 1) * (q[0] + 2) // 2 * ((q[1] + 1) * (q[1] + 2) // 2)
                ns = na_ // ngrp
                for gi in range(ngrp):
                    body.append(gen_grad_class_z(*q, list(range(gi * ns, (gi + 1) * ns)), "grad_%s_g%d" % (tag, gi)))
                calls = ["grad_%s_g%d<T, BOYS, MODE>(bra, bq, nbp, kr, nkp, bt, G, qtau, g);" % (tag, gi) for gi in range(ngrp)]
            body.append("template <int LA, int LB, int LC, int LD, typename T, int BOYS, int MODE, int GRP, typename KR>")
            body.append("__device__ __forceinline__ void grad_class(const double* __restrict__ bra, const double* __restrict__ bq, int nbp,")
            body.append("    const KR& kr, int nkp, const BoysTables& bt, const T* __restrict__ G, double qtau, T* __restrict__ g)")
            body.append("{")
            for gi, c in enumerate(calls):
                body.append("  %sif constexpr (GRP == %d) %s" % ("" if gi == 0 else "else ", gi, c))
            body.append("}")
        body.append("}  // namespace")
        body.append('#include "hfk_eri_impl.hpp"')
        if grad_coop:
            body.append('#include "hfk_coop_impl.hpp"')
            body.append("int hfk_dispatch_grad_%s(int boys, int mode, int, const QuartetArgs& a, hipStream_t s)" % tag)
            body.append("{ return dispatch_coop<%d, %d, %d, %d>(HFK_OP_GRAD, 0, boys, mode, a, s); }" % q)
        else:
            body.append('#include "hfk_grad_impl.hpp"')
            body.append("int hfk_dispatch_grad_%s(int boys, int mode, int prec, const QuartetArgs& a, hipStream_t s)" % tag)
            body.append("{ return dispatch_grad<%d, %d, %d, %d>(boys, mode, prec, a, s); }" % q)
        body.append("int hfk_dispatch_%s(int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)" % tag)
        body.append("{ return dispatch<%d, %d, %d, %d>(op, prec, boys, mode, a, s); }" % q)
        path = os.path.join(outdir, "eri_%s.hip" % tag)
        txt = "\n".join(body) + "\n"
        if not os.path.exists(path) or open(path).read() != txt:
            with open(path, "w") as f:
                f.write(txt)
    d = ["// Generated by gen_eri.py -- do not edit.", "#include <hip/hip_runtime.h>", '#include "hfk_internal.hpp"']
    for q in qs:
        d.append("int hfk_dispatch_%d%d%d%d(int, int, int, int, const QuartetArgs&, hipStream_t);" % q)
    for q in qs:
        d.append("int hfk_dispatch_grad_%d%d%d%d(int, int, int, const QuartetArgs&, hipStream_t);" % q)
    d.append("int hfk_launch_grad(int la, int lb, int lc, int ld, int boys, int mode, int prec, const QuartetArgs& a, hipStream_t s)")
    d.append("{")
    for q in qs:
        d.append("    if (la == %d && lb == %d && lc == %d && ld == %d) return hfk_dispatch_grad_%d%d%d%d(boys, mode, prec, a, s);" % (q + q))
    d.append("    return -2;")
    d.append("}")
    d.append("int hfk_launch_quartets(int la, int lb, int lc, int ld, int op, int prec, int boys, int mode, const QuartetArgs& a, hipStream_t s)")
    d.append("{")
    for q in qs:
        d.append("    if (la == %d && lb == %d && lc == %d && ld == %d) return hfk_dispatch_%d%d%d%d(op, prec, boys, mode, a, s);" % (q + q))
    d.append("    return -2;")
    d.append("}")
    path = os.path.join(outdir, "eri_dispatch.hip")
    txt = "\n".join(d) + "\n"
    if not os.path.exists(path) or open(path).read() != txt:
        with open(path, "w") as f:
            f.write(txt)
    print(" ".join("eri_%d%d%d%d" % q for q in qs) + " eri_dispatch")


if __name__ == "__main__":
    main()
