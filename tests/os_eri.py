"""Independent primitive ERI by Obara-Saika / Head-Gordon-Pople recursion
(VRR to [e0|f0]^(m), then HRR), with the Boys function from mpmath.  Used
only to pin the McMurchie-Davidson oracle; unnormalised primitives."""
import math
from functools import lru_cache

import mpmath


def boys(m, T):
    if T < 1e-15:
        return 1.0 / (2 * m + 1)
    return float(mpmath.gammainc(m + 0.5, 0, T) / (2 * mpmath.mpf(T) ** (m + 0.5)))


def prim_eri(A, la, a, B, lb, b, C, lc, c, D, ld, d):
    """(ab|cd) over x^l y^m z^n e^{-a r^2} primitives; la etc. are (lx,ly,lz)."""
    p, q = a + b, c + d
    P = [(a * A[i] + b * B[i]) / p for i in range(3)]
    Q = [(c * C[i] + d * D[i]) / q for i in range(3)]
    rho = p * q / (p + q)
    W = [(p * P[i] + q * Q[i]) / (p + q) for i in range(3)]
    AB2 = sum((A[i] - B[i]) ** 2 for i in range(3))
    CD2 = sum((C[i] - D[i]) ** 2 for i in range(3))
    PQ2 = sum((P[i] - Q[i]) ** 2 for i in range(3))
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(-a * b / p * AB2 - c * d / q * CD2)
    T = rho * PQ2
    Ltot = sum(la) + sum(lb) + sum(lc) + sum(ld)
    F = [boys(m, T) for m in range(Ltot + 1)]

    @lru_cache(maxsize=None)
    def vrr(e, f, m):
        # [e 0 | f 0]^(m), e and f are exponent tuples
        if min(e) < 0 or min(f) < 0:
            return 0.0
        if sum(e) == 0 and sum(f) == 0:
            return pref * F[m]
        if sum(e) > 0:
            i = next(k for k in range(3) if e[k] > 0)
            e1 = tuple(e[k] - (k == i) for k in range(3))
            e2 = tuple(e[k] - 2 * (k == i) for k in range(3))
            f1 = tuple(f[k] - (k == i) for k in range(3))
            v = (P[i] - A[i]) * vrr(e1, f, m) + (W[i] - P[i]) * vrr(e1, f, m + 1)
            if e1[i] > 0:
                v += e1[i] / (2 * p) * (vrr(e2, f, m) - rho / p * vrr(e2, f, m + 1))
            if f[i] > 0:
                v += f[i] / (2 * (p + q)) * vrr(e1, f1, m + 1)
            return v
        i = next(k for k in range(3) if f[k] > 0)
        f1 = tuple(f[k] - (k == i) for k in range(3))
        f2 = tuple(f[k] - 2 * (k == i) for k in range(3))
        v = (Q[i] - C[i]) * vrr(e, f1, m) + (W[i] - Q[i]) * vrr(e, f1, m + 1)
        if f1[i] > 0:
            v += f1[i] / (2 * q) * (vrr(e, f2, m) - rho / q * vrr(e, f2, m + 1))
        return v

    @lru_cache(maxsize=None)
    def hrr(ea, eb, ec, ed):
        # (ab|cd) from [e0|f0] by (a, b+1_i) = (a+1_i, b) + AB_i (a b)
        if sum(eb) > 0:
            i = next(k for k in range(3) if eb[k] > 0)
            eb1 = tuple(eb[k] - (k == i) for k in range(3))
            ea1 = tuple(ea[k] + (k == i) for k in range(3))
            return hrr(ea1, eb1, ec, ed) + (A[i] - B[i]) * hrr(ea, eb1, ec, ed)
        if sum(ed) > 0:
            i = next(k for k in range(3) if ed[k] > 0)
            ed1 = tuple(ed[k] - (k == i) for k in range(3))
            ec1 = tuple(ec[k] + (k == i) for k in range(3))
            return hrr(ea, eb, ec1, ed1) + (C[i] - D[i]) * hrr(ea, eb, ec, ed1)
        return vrr(ea, ec, 0)

    return hrr(tuple(la), tuple(lb), tuple(lc), tuple(ld))
